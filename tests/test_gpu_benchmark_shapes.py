"""GPU against the oracle at the benchmark's own shapes.

The 192 x 256 x 64-cell LES workload's cross-section, 192 (radial) x 256 (circumferential)
cells, with a 33-cell span, the workload's wall spacing (0.001 D), outer radius (25 D), Re 3900
and seeded start.  At these sizes the stencil kernels' 64-wide I tiles meet at seams (I = 64,
128), the K-marching kernels run more than one 16-plane segment, and the line scans hold three
(I lines of 192 rows) and four (J lines of 256 rows) rows per lane -- none of which the small
golden fixtures reach.

* Predictor, stage by stage through the entry points: NL before and after the SGS term, the
  dynamic coefficient, nu_t, SGS_x,y,z, DIF, S_u,v,w and u*, v*, w* against the oracle's
  stages (RLES/main.f90:42-88).
* One full time step (the benchmark's fused step, fvm_step) with the mixed Poisson form driven
  to RESIP <= 1e-10: the oracle evaluates the reference's scaled residual (iter_p.f90:45-59)
  of the GPU's pressure on its own predictor, then projects with that pressure
  (main.f90:105-150); velocity, pressure and the force coefficients must agree.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NI, NJ, NK = 193, 257, 34
RE, DT = 3900.0, 2e-4   # the benchmark's start-up dt (dt / 10)
I = (slice(1, NI), slice(1, NJ), slice(1, NK))


def rel(a, b, sl=I):
    a, b = a[sl], b[sl]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def case():
    from fvm_amd.grid import make_case
    (xg, yg, zg), ic = make_case(NI, NJ, NK, h1=0.001, r_out=25.0, sigma=0.05)
    state = [ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")]
    return xg, yg, zg, state


def oracle(case):
    from oracle.oracle3d import Oracle3D
    xg, yg, zg, state = case
    o = Oracle3D(NI, NJ, NK, RE, DT, math.pi, xg, yg, zg)
    o.set_state(*state)
    return o


def test_predictor_stages_at_benchmark_cross_section(case):
    from fvm_amd.solver import Solver3D
    xg, yg, zg, state = case
    print(f"predictor stages {NI - 1}x{NJ - 1}x{NK - 1} cells", flush=True)
    s = Solver3D(NI, NJ, NK, RE, DT, xg, yg, zg, use_les=True, nswpp=50, resimax=1e-10,
                 poisson_mode=3)
    s.set_state(*state)
    o = oracle(case)
    s.begin_step()
    s.contra_vel()
    s.nl()
    o.pred_nl()
    for c in ("nlx", "nly", "nlz"):
        assert rel(s.get(c), o.array(c)) < 1e-14, c
    c_avg = s.stress_tensor_dy()
    o.pred_sgs()
    assert abs(c_avg - o.scalar("c_avg")) <= 1e-12 * abs(o.scalar("c_avg")), c_avg
    assert rel(s.get("nut"), o.array("nut")) < 1e-12
    for c in ("sgsx", "sgsy", "sgsz", "nlx", "nly", "nlz"):
        assert rel(s.get(c), o.array(c)) < 1e-12, c
    s.dif()
    s.p_source()
    o.pred_dif_ps()
    for c in ("difx", "dify", "difz", "s_u", "s_v", "s_w"):
        assert rel(s.get(c), o.array(c)) < 1e-14, c
    s.iter_u()
    o.pred_iter_u()
    for c in ("u", "v", "w"):
        assert rel(s.get(c), o.array(c)) < 1e-12, c


@pytest.mark.parametrize("mode", [3, 2])
def test_converged_step_at_benchmark_cross_section(case, mode):
    from fvm_amd.solver import Solver3D
    xg, yg, zg, state = case
    print(f"converged step mode {mode}", flush=True)
    s = Solver3D(NI, NJ, NK, RE, DT, xg, yg, zg, use_les=True, nswpp=50, resimax=1e-10,
                 poisson_mode=mode)
    s.set_state(*state)
    info = s.step()
    assert info["status"] == 0 and info["resip"] <= 1e-10, info
    o = oracle(case)
    o.predictor(True)
    o.correct_predictor()
    # the GPU's pressure before its reference shift, ghosts included
    o.array("p")[...] = s.get("p") + info["pref"]
    res = o.p_residual()
    print(f"GPU resip {info['resip']:.3e}, oracle residual of the GPU pressure {res:.3e}",
          flush=True)
    assert res <= 1e-9, res
    o.finish_step()
    for c in ("u", "v", "w"):
        assert rel(s.get(c), o.array(c), (slice(None),) * 3) < 1e-10, c
    assert rel(s.get("p"), o.array("p"), (slice(None),) * 3) < 1e-10
    assert abs(info["pref"] - o.scalar("pref")) <= 1e-10 * max(abs(o.scalar("pref")), 1.0)
    for k, ok in (("cd", "cd"), ("cl", "cl"), ("cd_p", "cdp"), ("cd_v", "cdv")):
        assert abs(info[k] - o.scalar(ok)) <= 1e-9 * max(abs(o.scalar(ok)), 1.0), k
    assert abs(info["cfl"] - o.scalar("cfl")) <= 1e-10 * o.scalar("cfl")
