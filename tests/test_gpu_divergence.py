"""fvm_divergence: the discrete divergence of the face volume fluxes of the current velocity
(CONTRA_VEL's UE..WB, RLES/contra_vel.f90:13-56, summed as ITER_P's source, iter_p.f90:31-34),
against a numpy restatement from the fields and metrics the context holds."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def np_divergence(s):
    g = {k: s.get(k) for k in ("u", "v", "w", "djac", "xix", "xiy", "xiz", "etx", "ety", "etz",
                               "ztx", "zty", "ztz")}
    U, V, W, J = g["u"], g["v"], g["w"], g["djac"]
    ju = [J * (g["xix"] * U + g["xiy"] * V + g["xiz"] * W),
          J * (g["etx"] * U + g["ety"] * V + g["etz"] * W),
          J * (g["ztx"] * U + g["zty"] * V + g["ztz"] * W)]
    ni, nj, nk = (n - 1 for n in U.shape)
    for a in ju:  # periodic ghosts in J and K (MPI_X collapsed to one rank)
        a[:, 0, :] = a[:, nj - 1, :]
        a[:, nj, :] = a[:, 1, :]
        a[:, :, 0] = a[:, :, nk - 1]
        a[:, :, nk] = a[:, :, 1]
    I, Jc, K = slice(1, ni), slice(1, nj), slice(1, nk)
    c = (I, Jc, K)
    ue = 0.5 * (ju[0][c] + ju[0][2:ni + 1, Jc, K])
    uw = 0.5 * (ju[0][c] + ju[0][0:ni - 1, Jc, K])
    uw[0, :, :] = 0.0  # UW(I = 1) = 0 (contra_vel.f90:56)
    vn = 0.5 * (ju[1][c] + ju[1][I, 2:nj + 1, K])
    vs = 0.5 * (ju[1][c] + ju[1][I, 0:nj - 1, K])
    wt = 0.5 * (ju[2][c] + ju[2][I, Jc, 2:nk + 1])
    wb = 0.5 * (ju[2][c] + ju[2][I, Jc, 0:nk - 1])
    dv = ue - uw + vn - vs + wt - wb
    a = np.abs(dv)
    return {"max": float(a.max()), "mean": float(a.mean()),
            "rel": float(a.sum() / (np.abs(ue) + np.abs(vn) + np.abs(wt)).sum())}


def test_divergence_matches_numpy_after_converged_steps():
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    ni, nj, nk = 33, 65, 9
    (xg, yg, zg), ic = make_case(ni, nj, nk, h1=0.002, r_out=25.0, sigma=0.05, seed=11)
    s = Solver3D(ni, nj, nk, 3900.0, 2e-4, xg, yg, zg, ref_a=math.pi, use_les=True, nswpp=50,
                 resimax=1e-10, poisson_mode=2)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    for _ in range(2):
        info = s.step()
        assert info["resip"] <= 1e-10, info
    gpu = s.divergence()
    ref = np_divergence(s)
    print("divergence GPU", gpu, "numpy", ref, flush=True)
    assert gpu["max"] == pytest.approx(ref["max"], rel=1e-12, abs=0)
    for k in ("mean", "rel"):
        assert gpu[k] == pytest.approx(ref[k], rel=1e-10, abs=0), k
    # the start's seeded noise is not divergence-free; the projected velocity's face fluxes are
    # much closer to it (the collocated projection leaves an O(h^2) remainder)
    s0 = Solver3D(ni, nj, nk, 3900.0, 2e-4, xg, yg, zg, ref_a=math.pi, use_les=True, nswpp=50,
                  resimax=1e-10, poisson_mode=2)
    s0.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    d0 = s0.divergence()
    print("seeded start", d0, flush=True)
    assert gpu["rel"] < d0["rel"]
