"""GPU parity of the passive temperature (NL(T_N,ADV), the subgrid heat flux, DIF(T_N,DIFT),
ITER_T + BCT and the Nusselt number) against the reference's dumps of two converged steps.

Tolerances: the temperature path shares the velocity path's kernels; the dynamic coefficient
of the heat flux is a tree reduction and the test filter's terms are grouped by plane
(1e-12 relative for C_t, the eddy conductivity and SGS_T); the line solves reproduce the
reference's sweeps up to rounding (1e-10 for T after the step), the Nusselt number is a tree
sum over the wall (1e-10).
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def rel(a, b, sl=(slice(None),) * 3):
    a, b = a[sl], b[sl]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make_gpu(g, **kw):
    from fvm_amd.solver import Solver3D
    s = Solver3D(int(g["ni"]), int(g["nj"]), int(g["nk"]), float(g["re"]), float(g["dt"]),
                 g["xg"], g["yg"], g["zg"], ref_a=float(g["ref_a"]), use_les=True, **kw)
    s.set_state(*(g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    s.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    return s


@pytest.mark.parametrize("mode", [0, 2])
def test_thermal_steps_match_reference(mode):
    g = load_golden("rles_9x21x5_les_thermal")
    s = make_gpu(g, nswpp=200000, resimax=1e-13, poisson_mode=mode, check_every=4)
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    I = (slice(1, ni), slice(1, nj), slice(1, nk))
    for st in (1, 2):
        info = s.step()
        assert info["status"] == 0 and info["resip"] <= 1e-13
        assert abs(np.sqrt(info["c_t"]) - g["ct_sqrt"][st - 1]) <= 1e-12 * g["ct_sqrt"][st - 1]
        assert rel(s.get("kappa"), g[f"s{st}_kappa"]) < 1e-10, st
        assert rel(s.get("sgst"), g[f"s{st}_sgst"], I) < 1e-10, st
        assert rel(s.get("adv"), g[f"s{st}_adv"], I) < 1e-10, st
        assert rel(s.get("dift"), g[f"s{st}_dift"], I) < 1e-10, st
        assert rel(s.get("t"), g[f"s{st}_t"]) < 1e-10, st
        for c in ("u", "v", "w"):
            assert rel(s.get(c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        nu = g["scalars"][st - 1][10]
        assert abs(info["nu"] - nu) <= 1e-10 * abs(nu), (info["nu"], nu)
    s.close()


def test_thermal_first_step_stages():
    """Step 1 stage by stage: gradient, filtered temperature and the SGS heat flux."""
    g = load_golden("rles_9x21x5_les_thermal")
    s = make_gpu(g, nswpp=200000, resimax=1e-13)
    s.begin_step()
    s.contra_vel()
    s.nl()
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    I = (slice(1, ni), slice(1, nj), slice(1, nk))
    assert rel(s.get("adv"), g["s1_adv0"], I) < 1e-14
    s.stress_tensor_dy()
    for t in ("dtdx", "dtdy", "dtdz", "that"):
        assert rel(s.get(t), g["s1_" + t]) < 1e-13, t
    assert rel(s.get("sgst"), g["s1_sgst"], I) < 1e-11
    s.close()
