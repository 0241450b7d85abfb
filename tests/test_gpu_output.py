"""GPU parity of the field output of WRITEDATA (|omega| of CALC_VOR, lambda_2 by FIND_ROOT with
its ghost-face extrapolation, nodal averages) on the reference's own state after two thermal
steps, and the wall y+ of CALCCOEF.  |omega| and the nodal averages use only +, -, *, sqrt in
the reference's order (bit-exact); lambda_2 goes through acos, cos, sin and pow, whose device implementations differ
from the host's in the last bits (1e-12 relative to the field maximum)."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def test_writedata_fields_match_reference():
    from fvm_amd.solver import Solver3D
    g = load_golden("rles_9x21x5_les_thermal")
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    s = Solver3D(ni, nj, nk, float(g["re"]), float(g["dt"]), g["xg"], g["yg"], g["zg"],
                 ref_a=float(g["ref_a"]), use_les=True)
    s.set_state(g["s2_u"], g["s2_v"], g["s2_w"], g["s2_pref"], g["in_nlx"], g["in_nly"],
                g["in_nlz"])
    s.set_temperature(g["s2_t"], g["in_adv"], float(g["pe"]))
    print("field output on the 8x20x4 thermal case", flush=True)
    s.writedata_fields()
    assert np.array_equal(s.get("vorm"), g["out_vorm"])
    lam = s.get("lamda2")
    assert np.abs(lam - g["out_lamda2"]).max() <= 1e-12 * np.abs(g["out_lamda2"]).max()
    for a, t in (("pg", "out_pg"), ("ug", "out_ug"), ("vg", "out_vg"), ("wg", "out_wg"),
                 ("tg", "out_tg")):
        assert np.array_equal(s.get(a)[1:, 1:, 1:], g[t]), a
    assert np.array_equal(s.get("vorg")[1:, 1:, 1:], g["out_dum3d"][1:, 1:, 1:])
    lg = s.get("lamdag")[1:, 1:, 1:]
    assert np.abs(lg - g["out_lamdag"]).max() <= 1e-12 * np.abs(g["out_lamdag"]).max()
    # wall y+ of the last CALCCOEF: +, *, /, sqrt of the wall table in the reference's order
    yp = s.get("yplus")[1, 1:nj, 1:nk]
    assert np.abs(yp - g["out_yplus"]).max() <= 1e-14 * np.abs(g["out_yplus"]).max()
    s.close()
