"""Time averages of POST_PROCESS on the GPU against the reference (fixture
rles_9x21x5_les_post: one converged thermal LES step at dt = 4e-3, then three steps of the
averaging phase at DTIME_POST = 1e-3 and the finalization).  The running sums, the in-place
means and Reynolds stresses, the node averages and the wall Cf / Cp / Nu are compared at the
FP64 field tolerance 1e-10 (relative to the field's largest value); the Tecplot values the
reference printed (9 digits after the point) at 1e-8."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

SUMS = ("u", "v", "w", "p", "t", "uu", "uv", "uw", "vv", "vw", "ww", "sgs11", "sgs12", "sgs13",
        "sgs22", "sgs23", "sgs33")


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_post_process_matches_reference():
    from fvm_amd.solver import Solver3D
    g = load_golden("rles_9x21x5_les_post")
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    s = Solver3D(ni, nj, nk, float(g["re"]), float(g["dt"]), g["xg"], g["yg"], g["zg"],
                 ref_a=float(g["ref_a"]), use_les=True, nswpp=200000, resimax=1e-13,
                 check_every=4)
    s.set_state(*(g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    s.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    steps, npost = int(g["scalars"].shape[0]), int(g["npost"])
    for st in range(1, steps + 1):
        if st == steps - npost + 1:
            s.post_begin()
            s.set_dtime(float(g["dt_post"]))
        info = s.step()
        assert info["status"] == 0 and info["resip"] <= 1e-13
    for c, key in (("u", "u"), ("v", "v"), ("w", "w"), ("p", "pref"), ("t", "t")):
        assert rel(s.get(c), g[f"s{steps}_{key}"]) < 1e-10, c
    for q in SUMS:
        assert rel(s.get("sum_" + q), g["post_sum_" + q]) < 1e-10, q
    assert s.post_finalize() == int(g["niter_pos"]) == npost
    for q in ("u", "v", "w", "p", "t", "uu", "uv", "uw", "vv", "vw", "ww"):
        assert rel(s.get("sum_" + q), g["post_fin_" + q]) < 1e-10, q
    n = (slice(1, None), slice(1, None), slice(1, None))
    for q, name in (("pg", "avg_pg"), ("tg", "avg_tg"), ("uu", "rs_uu"), ("uv", "rs_uv"),
                    ("uw", "rs_uw"), ("vv", "rs_vv"), ("vw", "rs_vw"), ("ww", "rs_ww")):
        assert rel(s.get(name)[n], g["post_node_" + q]) < 1e-10, q
    for q, name in (("pg", "avg_pg"), ("ug", "avg_ug"), ("vg", "avg_vg"), ("wg", "avg_wg"),
                    ("tg", "avg_tg"), ("urms", "rms_ug"), ("vrms", "rms_vg"),
                    ("wrms", "rms_wg"), ("uu", "rs_uu"), ("uv", "rs_uv"), ("uw", "rs_uw"),
                    ("vv", "rs_vv"), ("vw", "rs_vw"), ("ww", "rs_ww")):
        assert rel(s.get(name)[n], g["plt_" + q]) < 1e-8, q
    w = (1, slice(1, nj), slice(1, nk))
    for q in ("cf", "cp", "nu"):
        assert rel(s.get("surf_" + q)[w], g["post_surf_" + q]) < 1e-10, q
        assert rel(s.get("surf_" + q)[w], g["plt_surf_" + q]) < 1e-8, q
    # the accumulation has stopped: a further step leaves the results alone
    before = s.get("rs_uu")
    s.step()
    assert np.array_equal(s.get("rs_uu"), before)
