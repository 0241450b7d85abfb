"""The AF predictor on a grid whose span holds more than 64 cells.

Spans of up to 64 cells keep a whole K line of the tridiagonal solve in registers; longer spans
go through the general K-line Thomas solve, which streams blocks of rows from memory.  The
golden fixtures all have short spans, so this case checks the general solve against the CPU
oracle on a seeded 8 x 12 x 80-cell O-grid, stage by stage up to the three ITER_U solves.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NI, NJ, NK = 9, 13, 81


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_predictor_long_span_matches_oracle():
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    from oracle.oracle3d import Oracle3D
    (xg, yg, zg), ic = make_case(NI, NJ, NK, h1=0.01, r_out=10.0, sigma=0.05)
    state = [ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")]
    o = Oracle3D(NI, NJ, NK, 3900.0, 0.002, np.pi, xg, yg, zg)
    o.set_state(*state)
    o.predictor(True)
    s = Solver3D(NI, NJ, NK, 3900.0, 0.002, xg, yg, zg, use_les=True)
    s.set_state(*state)
    s.begin_step()
    s.contra_vel()
    s.nl()
    s.stress_tensor_dy()
    s.dif()
    s.p_source()
    s.iter_u()
    I = (slice(1, NI), slice(1, NJ), slice(1, NK))
    for c in ("u", "v", "w"):
        assert rel(s.get(c)[I], o.array(c)[I]) < 1e-12, c
