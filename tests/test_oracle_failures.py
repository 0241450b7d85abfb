"""The oracle's ITER_P on a zero right-hand side with p = 0: the scaled residual is 0/0 = NaN,
for which the reference's test 'RESIP .GT. RESIMAX' is false, so its sweep loop ends after one
sweep with p unchanged (iter_p.f90:36-62).  The GPU maps that state to RESIP = 0 (converged)."""
import math

import numpy as np

from oracle.oracle3d import Oracle3D


def test_zero_rhs_ends_after_one_sweep():
    from fvm_amd.grid import make_case
    ni, nj, nk = 9, 13, 5
    (xg, yg, zg), ic = make_case(ni, nj, nk, h1=0.02, r_out=6.0, sigma=0.0)
    o = Oracle3D(ni, nj, nk, 3900.0, 2e-3, math.pi, xg, yg, zg)
    z = np.zeros((ni + 1, nj + 1, nk + 1), order="F")
    o.set_state(z, z, z, z, z, z, z)
    o.L.orc3d_contra_vel(o.h)
    o.L.orc3d_iter_p(o.h, 20, 1e-8)
    assert o.scalar("niter") == 1
    assert math.isnan(o.scalar("resip"))
    assert not np.any(o.array("p"))
