"""The FP32 change form of the 3D AF I and J stages (af_mode 2) on a ragged grid.

70 x 96 x 10 interior cells: I rows not a multiple of the 64-cell tiles of the J-line unpack
and of the I-stage add, K planes not a multiple of the add's four planes per thread, lines of
two rows per lane (the vector row loads of the J-line pass).  From a seeded start, the predictor
(CONTRA_VEL to ITER_U, RLES/main.f90:48-88) of the change form equals the FP64
sweeps' to the FP32 resolution of the change the stages make.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(af_mode):
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    ni, nj, nk = 71, 97, 11
    (xg, yg, zg), ic = make_case(ni, nj, nk, h1=0.002, r_out=25.0, sigma=0.05, seed=23)
    s = Solver3D(ni, nj, nk, 3900.0, 2e-3, xg, yg, zg, use_les=True, nswpp=50, resimax=1e-10,
                 poisson_mode=2, af_mode=af_mode)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    return s


def predictor(s):
    s.begin_step()
    s.contra_vel()
    s.nl()
    s.stress_tensor_dy()
    s.dif()
    s.p_source()
    s.iter_u()
    return {c: s.get(c) for c in ("u", "v", "w")}


def test_fp32_change_form_on_ragged_grid():
    s0 = make(0)
    start = {c: s0.get(c) for c in ("u", "v", "w")}
    a = predictor(s0)
    b = predictor(make(2))
    for c in ("u", "v", "w"):
        ch = np.abs(a[c] - start[c])[1:-1, 1:-1, 1:-1].max()
        e = float(np.abs(b[c] - a[c])[1:-1, 1:-1, 1:-1].max() / ch)
        print(f"ragged predictor {c}: |FP32 change form - FP64 sweeps| / |change| {e:.2e}",
              flush=True)
        assert e < 2e-6, c
