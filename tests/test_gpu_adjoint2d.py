"""Discrete adjoint of the batched 2D step (fvm_batch2d_adjoint_*, fvm_amd.adjoint).

* The differentiable cylinder case: 97 radial x 193 circumferential nodes (wall spacing 0.002 D,
  outer radius 30 D), Re 150, dt 0.005, FP64, seeded start; 128 wall-normal blowing / suction
  velocities drawn N(0, 0.01); objective the mean C_D over the last 400 of 600 steps.  The
  adjoint gradient (sqrt(N) checkpointing) against central finite differences with h = 1e-6 on
  8 randomly chosen controls; the gradient from a mixed-precision forward trajectory (FP32
  Poisson transforms inside the FP64 defect correction) against the FP64 one; the
  torch.autograd path against the direct call.
* A short horizon on a small grid: the whole 128-component gradient against central
  differences, and the adjoint state's own consistency (the adjoint Poisson solve converges).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6


def _problem(ni, nj, re, dt, nsteps, navg, mixed=False, seed=11):
    from fvm_amd.adjoint import DragProblem
    from fvm_amd.batch2d import Batch2D
    from fvm_amd.grid import make_case
    (xg, yg, _), ic = make_case(ni, nj, 1, h1=0.002, r_out=30.0, sigma=1e-3, seed=seed)
    b = Batch2D([dict(ni=ni, nj=nj, re=re, dt=dt, xg=xg[:, :, 0], yg=yg[:, :, 0])],
                mixed=mixed, resimax=1e-14, npass=6)
    x0 = [tuple(ic[k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly"))]
    return DragProblem(b, x0, nsteps, navg, tol=1e-13, maxpass=60)


def _controls(seed=5):
    return np.random.default_rng(seed).normal(0.0, 0.01, (1, 128))


def _fd(pr, c, k):
    cp, cm = c.copy(), c.copy()
    cp[0, k] += H
    cm[0, k] -= H
    return (pr.run(cp, keep=False)[0] - pr.run(cm, keep=False)[0]) / (2 * H)


def test_gradient_against_central_differences_re150():
    pr = _problem(97, 193, 150.0, 0.005, 600, 400)
    c = _controls()
    j = pr.run(c)[0]
    g = pr.gradient(c)[0]
    print(f"mean C_D {j:.12f}, |g| {np.abs(g).max():.3e}, adjoint Poisson residual "
          f"{pr.worst_ares:.1e}, forward RESIP max {pr.max_resip:.1e}", flush=True)
    assert np.isfinite(g).all() and pr.worst_ares <= 1e-13
    ks = np.random.default_rng(7).choice(128, 8, replace=False)
    gmax = np.abs(g).max()
    for k in ks:
        fd = _fd(pr, c, k)
        print(f"control {k:3d}: adjoint {g[k]: .10e}  FD {fd: .10e}  diff {g[k] - fd: .2e}",
              flush=True)
        assert abs(g[k] - fd) <= 1e-5 * gmax + 1e-9, (k, g[k], fd)


def test_mixed_precision_trajectory_gradient_against_fp64():
    c = _controls()
    g64 = _problem(97, 193, 150.0, 0.005, 600, 400).gradient(c)[0]
    pm = _problem(97, 193, 150.0, 0.005, 600, 400, mixed=True)
    gm = pm.gradient(c)[0]
    err = np.abs(gm - g64).max() / np.abs(g64).max()
    print(f"mixed-precision trajectory gradient vs FP64: {err:.2e} (forward RESIP max "
          f"{pm.max_resip:.1e})", flush=True)
    assert err <= 1e-6


def test_torch_autograd_matches_direct_gradient():
    import torch
    from fvm_amd.adjoint import mean_drag
    pr = _problem(33, 65, 150.0, 0.005, 40, 20)
    c = _controls()
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    j = mean_drag(pr, ct)
    (3.0 * j.sum()).backward()
    g = pr.gradient(c)
    assert np.allclose(ct.grad.numpy(), 3.0 * g, rtol=1e-12, atol=0.0)


def test_full_gradient_short_horizon():
    pr = _problem(33, 129, 100.0, 0.005, 12, 6)
    c = _controls()
    pr.run(c)
    g = pr.gradient(c)[0]
    fd = np.array([_fd(pr, c, k) for k in range(128)])
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"short horizon: |g| {np.abs(g).max():.3e}, max rel diff to FD {err:.2e}", flush=True)
    assert err <= 1e-5
