"""GPU parity of the 2D solver (fvm_create2d) against the flang-built 2D reference
(cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/, golden fixture): three steps with the
pressure equation converged (reference: scaled residual 1e-15); velocities and pressure within 1e-10 of the field maximum, the
2D CALCCOEF coefficients (C_D, C_L and their pressure / viscous parts, C_pb, CFL) within 1e-9."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("mode", [2, 3])
def test_2d_steps_match_reference(mode):
    from fvm_amd.solver import Solver2D
    g = load_golden("r2d_17x33_re100_converged")
    ni, nj = int(g["ni"]), int(g["nj"])
    s = Solver2D(ni, nj, float(g["re"]), float(g["dt"]), g["xg"], g["yg"], nswpp=50,
                 resimax=1e-14, poisson_mode=mode)
    s.set_state(g["in_u"], g["in_v"], g["in_p"], g["in_nlx"], g["in_nly"])
    for st in (1, 2, 3):
        info = s.step()
        assert info["status"] == 0 and info["resip"] <= 1e-14, info
        for c in ("u", "v", "p"):
            assert rel(s.get(c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        row = g["scalars"][st - 1]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        for key, col in (("cfl", 3), ("cpb", 4), ("cd", 5), ("cl", 6), ("cd_p", 7),
                         ("cd_v", 8), ("cl_p", 9), ("cl_v", 10)):
            ref = row[col]
            assert abs(info[key] - ref) <= 1e-9 * max(abs(ref), 1.0), (st, key, info[key], ref)
