"""GPU parity of the 3D time step against the reference (golden dumps) and the CPU oracle.

Tolerances: the explicit stencils, the set-up metrics and the boundary conditions evaluate the
reference's expressions in the reference's order and are compared at 1e-13 relative (the
dynamic-Smagorinsky coefficient is a tree reduction, hence not bit-identical).  Line solves use
parallel cyclic reduction instead of the Thomas recursion (1e-12).  The Poisson equation is
relaxed in zebra order instead of the reference's lexicographic order, so pressure and the
projected velocity are compared after both have converged to the same fixed point
(rtol 1e-10 against the field maximum).
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
                 g["xg"], g["yg"], g["zg"], ref_a=float(g["ref_a"]), use_les=bool(g["use_les"]),
                 **kw)
    s.set_state(*(g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    return s


def inner(g):
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    return (slice(1, ni), slice(1, nj), slice(1, nk))


@pytest.mark.parametrize("name", ["rles_13x17x9_les_nswpp20", "rles_9x21x5_les_converged"])
def test_setup_matches_reference(name):
    g = load_golden(name)
    s = make_gpu(g)
    for t in ("x", "y", "z", "djac", "xix", "xiy", "xiz", "etx", "ety", "etz", "ztx", "zty",
              "ztz"):
        assert np.array_equal(s.get(t), g["setup_" + t]), t
    th = s.get("theta")
    assert rel(th, g["setup_theta"]) < 1e-15


def test_stages_match_reference():
    """Step 1 of the 20-sweep fixture, stage by stage up to the predictor correction."""
    g = load_golden("rles_13x17x9_les_nswpp20")
    s = make_gpu(g, nswpp=20, resimax=1e-4)
    I = inner(g)
    s.begin_step()
    s.contra_vel()
    s.nl()
    for c, t in (("nlx", "nl0x"),):
        assert rel(s.get(c), g["s1_" + t], I) < 1e-14, t
    c_avg = s.stress_tensor_dy()
    assert np.isfinite(c_avg)
    assert rel(s.get("nut"), g["s1_nut"]) < 1e-12
    for c in ("sgsx", "sgsy", "sgsz"):
        assert rel(s.get(c), g["s1_" + c], I) < 1e-12, c
    for c in ("nlx", "nly", "nlz"):
        assert rel(s.get(c), g["s1_" + c], I) < 1e-12, c
    s.dif()
    for c in ("difx", "dify", "difz"):
        assert np.array_equal(s.get(c)[I], g["s1_" + c][I]), c
    s.p_source()
    assert np.array_equal(s.get("s_u")[I], g["s1_su0"][I])
    s.iter_u()
    for c, t in (("u", "ustar"), ("v", "vstar"), ("w", "wstar")):
        assert rel(s.get(c)[I], g["s1_" + t][I]) < 1e-12, c
    s.correct(-1.0)
    assert rel(s.get("u"), g["s1_uhat"]) < 1e-12


def test_converged_steps_match_reference():
    """Two full steps with the pressure equation converged (fixture solved to RESIP 1e-13)."""
    g = load_golden("rles_9x21x5_les_converged")
    s = make_gpu(g, nswpp=200000, resimax=1e-13, check_every=4)
    for st in (1, 2):
        info = s.step()
        assert info["status"] == 0 and info["resip"] <= 1e-13
        for c in ("u", "v", "w"):
            assert rel(s.get(c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        assert rel(s.get("p"), g[f"s{st}_pref"]) < 1e-10, (st, "p")
        # SGS_x,y,z and nu_t are formed from the step's C, Delta^2 and S_ij when first read
        if st == 1:
            I = inner(g)
            for c in ("sgsx", "sgsy", "sgsz"):
                assert rel(s.get(c), g[f"s1_{c}"], I) < 1e-10, c
            assert rel(s.get("nut"), g["s1_nut"]) < 1e-10
        row = g["scalars"][st - 1]
        assert abs(info["cd"] - row[6]) <= 1e-9 * abs(row[6])
        assert abs(info["cl"] - row[7]) <= 1e-9 * max(abs(row[7]), 1.0)
        assert abs(info["cfl"] - row[4]) <= 1e-10 * row[4]


def test_mixed_precision_reaches_same_fixed_point():
    """FP32 sweeps inside FP64 defect correction converge to the FP64 fixed point."""
    g = load_golden("rles_9x21x5_les_converged")
    s = make_gpu(g, nswpp=200000, resimax=1e-13, poisson_mode=1, check_every=8)
    info = s.step()
    assert info["status"] == 0 and info["resip"] <= 1e-13
    for c in ("u", "v", "w"):
        assert rel(s.get(c), g[f"s1_{c}"]) < 1e-10, c
    assert rel(s.get("p"), g["s1_pref"]) < 1e-10


@pytest.mark.parametrize("mode", [2, 3])
def test_separable_defect_correction_same_fixed_point(mode):
    """Defect correction with the separable approximate inverse (FP64: mode 2, FP32: mode 3)
    converges to the reference's fixed point."""
    g = load_golden("rles_9x21x5_les_converged")
    s = make_gpu(g, nswpp=50, resimax=1e-13, poisson_mode=mode)
    for st in (1, 2):
        info = s.step()
        assert info["status"] == 0 and info["resip"] <= 1e-13, info
        assert info["niter"] <= (4 if mode == 2 else 8), info["niter"]
        for c in ("u", "v", "w"):
            assert rel(s.get(c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        assert rel(s.get("p"), g[f"s{st}_pref"]) < 1e-10, (st, "p")
