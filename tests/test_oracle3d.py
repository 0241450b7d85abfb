"""Pins the CPU oracle (oracle/fvm_oracle3d.c) to the reference solver and to closed forms.

The golden fixtures under tests/golden/ were written by the reference's own sources built with
flang (oracle/ref/build_ref.sh + oracle/ref/harness3d.f90, script tests/golden/make_golden.py).
The oracle follows the reference's loop order and expression association and must reproduce
every dumped stage bit for bit.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle.oracle3d import Oracle3D

FIXTURES = ["rles_13x17x9_les_nswpp20", "rles_9x21x5_les_converged"]


def make(g):
    o = Oracle3D(int(g["ni"]), int(g["nj"]), int(g["nk"]), float(g["re"]), float(g["dt"]),
                 float(g["ref_a"]), g["xg"], g["yg"], g["zg"])
    o.set_state(*(g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    return o


@pytest.mark.parametrize("name", FIXTURES)
def test_setup_bitexact(name):
    g = load_golden(name)
    o = make(g)
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    for t in ("x", "y", "z", "djac", "xix", "xiy", "xiz", "etx", "ety", "etz", "ztx", "zty",
              "ztz", "q12", "q13", "q23", "theta"):
        assert np.array_equal(o.array(t), g["setup_" + t]), t
    for t in ("de", "dn", "dt"):
        assert np.array_equal(o.array(t)[inner], g["setup_" + t][inner]), t


@pytest.mark.parametrize("name", FIXTURES)
def test_steps_bitexact(name):
    g = load_golden(name)
    o = make(g)
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    steps = int(g["scalars"].shape[0])
    for s in range(1, steps + 1):
        o.step(int(g["nswpp"]), float(g["resimax"]), bool(g["use_les"]))
        if f"s{s}_u" not in g:
            continue
        for t, a in (("nlx", "nlx"), ("nly", "nly"), ("nlz", "nlz"), ("difx", "difx"),
                     ("dify", "dify"), ("difz", "difz"), ("sgsx", "sgsx"), ("sgsy", "sgsy"),
                     ("sgsz", "sgsz")):
            assert np.array_equal(o.array(a)[inner], g[f"s{s}_{t}"][inner]), (s, t)
        assert np.array_equal(o.array("nut"), g[f"s{s}_nut"]), (s, "nut")
        for t in ("u", "v", "w"):
            assert np.array_equal(o.array(t), g[f"s{s}_{t}"]), (s, t)
        assert np.array_equal(o.array("p"), g[f"s{s}_pref"]), (s, "p")
        row = g["scalars"][s - 1]  # istep niter resip pref cfl cpb cd cl cd_p cd_v
        assert o.scalar("niter") == row[1]
        assert o.scalar("resip") == row[2]
        assert o.scalar("pref") == row[3]
        assert o.scalar("cfl") == row[4]
        assert o.scalar("cpb") == row[5]
        assert o.scalar("cd") == row[6] and o.scalar("cl") == row[7]
        assert o.scalar("cdp") == row[8] and o.scalar("cdv") == row[9]


def test_predictor_bitexact():
    """The predictor half of step 1 leaves the reference's u*, v*, w* (ITER_U outputs)."""
    g = load_golden("rles_13x17x9_les_nswpp20")
    o = make(g)
    o.predictor(bool(g["use_les"]))
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    for t, a in (("ustar", "u"), ("vstar", "v"), ("wstar", "w")):
        assert np.array_equal(o.array(a)[inner], g["s1_" + t][inner]), t


def cartesian(ni, nj, nk, h=(0.3, 0.7, 0.45)):
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    return (np.asfortranarray(1.0 + h[0] * i), np.asfortranarray(2.0 + h[1] * j),
            np.asfortranarray(h[2] * k))


def test_metric_closed_form_cartesian():
    """Away from the periodic J/K seams a Cartesian node set has exact central differences:
    J^-1 = hx hy hz, xi_x = 1/hx, eta_y = 1/hy, zeta_z = 1/hz, all cross terms zero."""
    h = (0.3, 0.7, 0.45)
    ni, nj, nk = 9, 11, 8
    xg, yg, zg = cartesian(ni, nj, nk, h)
    o = Oracle3D(ni, nj, nk, 100.0, 1e-3, 1.0, xg, yg, zg)
    s = (slice(1, ni), slice(2, nj - 2), slice(2, nk - 2))
    np.testing.assert_allclose(o.array("djac")[s], h[0] * h[1] * h[2], rtol=1e-13)
    np.testing.assert_allclose(o.array("xix")[s], 1 / h[0], rtol=1e-13)
    np.testing.assert_allclose(o.array("ety")[s], 1 / h[1], rtol=1e-13)
    np.testing.assert_allclose(o.array("ztz")[s], 1 / h[2], rtol=1e-13)
    for t in ("xiy", "xiz", "etx", "etz", "ztx", "zty"):
        assert np.abs(o.array(t)[s]).max() < 1e-12, t


def test_annulus_volume_and_wall_area():
    """Sum of J^-1 over the interior approximates the annulus volume pi (R^2 - r^2) Lz; the
    wall triangles sum to the polygonal lateral area."""
    from fvm_amd.grid import OGrid
    ni, nj, nk = 41, 129, 9
    g = OGrid(ni, nj, nk, 0.01, 2.0, np.pi)
    xg, yg, zg = g.nodes()
    o = Oracle3D(ni, nj, nk, 100.0, 1e-3, np.pi, xg, yg, zg)
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    vol = np.pi * (2.0 ** 2 - 0.5 ** 2) * np.pi
    assert abs(o.array("djac")[inner].sum() / vol - 1) < 2e-3
    area = o.array("area1")[1, 1:nj, 1:nk].sum() + o.array("area2")[1, 1:nj, 1:nk].sum()
    poly = (nj - 1) * 2 * 0.5 * np.sin(np.pi / (nj - 1)) * np.pi
    assert abs(area / poly - 1) < 1e-12


def test_operators_on_simple_fields():
    """DIF(const) = 0, NL(const) = const * sum of face fluxes, TEST_FILTER keeps constants and
    index-linear fields, P_SOURCE(x) = -J^-1 on a Cartesian box."""
    import ctypes
    h = (0.3, 0.7, 0.45)
    ni, nj, nk = 9, 11, 8
    xg, yg, zg = cartesian(ni, nj, nk, h)
    o = Oracle3D(ni, nj, nk, 100.0, 1e-3, 1.0, xg, yg, zg)
    L, H = o.L, o.h
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    shape = o.shape
    c = np.full(shape, 2.5, order="F")
    out = np.zeros(shape, order="F")
    L.orc3d_dif(H, ptr(c), ptr(out))
    assert np.abs(out).max() == 0.0
    rng = np.random.default_rng(0)
    for nm in ("u", "v", "w"):
        o.array(nm)[...] = rng.normal(size=shape)
    L.orc3d_contra_vel(H)
    L.orc3d_nl(H, ptr(c), ptr(out))
    div = (o.array("ue") - o.array("uw") + o.array("vnf") - o.array("vs") + o.array("wt")
           - o.array("wb"))
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    np.testing.assert_allclose(out[inner], 2.5 * div[inner], rtol=1e-12, atol=1e-12)
    L.orc3d_test_filter(H, ptr(c), ptr(out))
    np.testing.assert_allclose(out[inner], 2.5, rtol=1e-15)
    i, j, k = np.meshgrid(np.arange(ni + 1), np.arange(nj + 1), np.arange(nk + 1), indexing="ij")
    lin = np.asfortranarray(0.5 * i - 0.25 * j + 2.0 * k + 1.0)
    L.orc3d_test_filter(H, ptr(lin), ptr(out))
    np.testing.assert_allclose(out[inner], lin[inner], rtol=1e-14)
    px = np.asfortranarray(o.array("x").copy())
    L.orc3d_p_source(H, ptr(px))
    s = (slice(2, ni - 1), slice(2, nj - 2), slice(2, nk - 2))
    np.testing.assert_allclose(o.array("s_u")[s], -h[0] * h[1] * h[2], rtol=1e-12)
    assert np.abs(o.array("s_v")[s]).max() < 1e-13


def test_poisson_fixed_point_residual():
    """The converged golden pressure satisfies the full discrete equation (orthogonal part,
    explicit cross terms, wall Neumann fold, outflow AE=0, periodic ghosts) to the fixture's
    tolerance -- an ordering-independent property of the reference's iteration."""
    g = load_golden("rles_9x21x5_les_converged")
    o = make(g)
    o.step(int(g["nswpp"]), float(g["resimax"]), True)
    assert o.scalar("resip") <= 1e-13


def test_coldstart_step_bitexact_and_reference_divergence():
    """Wall spacing 0.001 D, dt 0.002, Re 3900, impulsive seeded start (the 192x256x64 LES
    case's parameters on a 32x64x8 grid): step 1 of the oracle equals the reference's bit for
    bit, and the reference's own CFL history grows by about 10x per step -- the impulsive start
    at this dt diverges in the reference scheme itself."""
    g = load_golden("rles_33x65x9_coldstart_dt2e-3")
    o = make(g)
    o.step(int(g["nswpp"]), float(g["resimax"]), bool(g["use_les"]))
    row = g["scalars"][0]
    assert o.scalar("niter") == row[1] and o.scalar("resip") == row[2]
    assert o.scalar("cfl") == row[4] and o.scalar("cd") == row[6]
    assert np.array_equal(o.array("u"), g["s1_u"])
    cfl = g["scalars"][:, 4]
    assert cfl[0] > 10 and cfl[1] > 5 * cfl[0] and cfl[2] > 50 * cfl[1]


def test_thermal_steps_bitexact():
    """The passive temperature (NL, SGS_T, DIF, ITER_T, BCT, Nusselt number) of two converged
    steps equals the reference's dumps bit for bit, and the velocity is unchanged by it."""
    g = load_golden("rles_9x21x5_les_thermal")
    o = make(g)
    o.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    inner = (slice(1, ni), slice(1, nj), slice(1, nk))
    for s in (1, 2):
        o.step(int(g["nswpp"]), float(g["resimax"]), True)
        for t in ("adv", "sgst", "dift"):
            assert np.array_equal(o.array(t)[inner], g[f"s{s}_{t}"][inner]), (s, t)
        for t in ("t", "kappa", "that", "dtdx", "dtdy", "dtdz"):
            assert np.array_equal(o.array(t), g[f"s{s}_{t}"]), (s, t)
        for t in ("u", "v", "w", "nut"):
            assert np.array_equal(o.array(t), g[f"s{s}_{t}"]), (s, t)
        row = g["scalars"][s - 1]  # ... cd_v, nu
        assert o.scalar("nu") == row[10]
        assert np.sqrt(o.scalar("c_t")) == g["ct_sqrt"][s - 1]
        assert np.sqrt(o.scalar("c_avg")) == g["cs_sqrt"][s - 1]


def test_thermal_closed_forms():
    """Without the dynamic model the temperature step is an affine map of T (the velocity is
    independent of T): the step of the mean of two temperature fields is the mean of their
    steps.  A cold fluid (T = TIN = 0) against the hot wall (T_WALL = 1) gives Nu > 0."""
    g = load_golden("rles_9x21x5_les_converged")
    rng = np.random.default_rng(7)
    out = []
    o = make(g)
    ta, tb = (rng.normal(0.0, 0.1, o.shape) for _ in range(2))
    for t0 in (ta, tb, 0.5 * (ta + tb)):
        o = make(g)
        o.set_temperature(np.asfortranarray(t0), np.zeros(o.shape, order="F"))
        o.step(int(g["nswpp"]), float(g["resimax"]), False)
        out.append(o.array("t").copy())
    mid = 0.5 * (out[0] + out[1])
    assert np.abs(out[2] - mid).max() <= 1e-13 * np.abs(mid).max()
    o2 = make(g)
    o2.set_temperature(np.zeros(o2.shape, order="F"), np.zeros(o2.shape, order="F"))
    o2.step(int(g["nswpp"]), float(g["resimax"]), True)
    assert o2.scalar("nu") > 0.0


def test_field_output_bitexact():
    """WRITEDATA's field output after the two thermal steps: |omega| (CALC_VOR), lambda_2 of
    S^2 + Omega^2 by FIND_ROOT with the ghost-face extrapolation, and the nodal averages of p,
    u, v, w, T, |omega|, lambda_2, equal to the reference's arrays bit for bit."""
    g = load_golden("rles_9x21x5_les_thermal")
    o = make(g)
    o.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    for _ in (1, 2):
        o.step(int(g["nswpp"]), float(g["resimax"]), True)
    o.writedata_fields()
    assert np.array_equal(o.array("lamda2"), g["out_lamda2"])
    assert np.array_equal(o.array("vorm"), g["out_vorm"])
    for a, t in (("pg", "out_pg"), ("ug", "out_ug"), ("vg", "out_vg"), ("wg", "out_wg"),
                 ("tg", "out_tg"), ("lamdag", "out_lamdag")):
        assert np.array_equal(o.array(a)[1:, 1:, 1:], g[t]), a
    assert np.array_equal(o.array("vorg")[1:, 1:, 1:], g["out_dum3d"][1:, 1:, 1:])
    nj, nk = int(g["nj"]), int(g["nk"])
    assert np.array_equal(o.array("yplus")[1, 1:nj, 1:nk], g["out_yplus"])


def test_yplus_geometry():
    """The wall y+ of CALCCOEF from vector geometry: for each wall triangle of the cell's wall
    face, n the unit normal of its plane, d_n the distance of the cell centre from that plane
    and u_t = u - (u.n) n the tangential velocity, y+ is the area-weighted sum over the two
    triangles of sqrt(Re |u_t| / d_n) d_n."""
    g = load_golden("rles_9x21x5_les_thermal")
    o = make(g)
    o.set_state(g["s2_u"], g["s2_v"], g["s2_w"], g["s2_pref"], g["in_nlx"], g["in_nly"],
                g["in_nlz"])
    o.calccoef()
    nj, nk, re = int(g["nj"]), int(g["nk"]), float(g["re"])
    P = np.stack([o.array(a)[1] for a in ("xg", "yg", "zg")], axis=-1)   # wall nodes (j, k)
    C = np.stack([o.array(a)[1] for a in ("x", "y", "z")], axis=-1)      # wall-cell centres
    U = np.stack([o.array(a)[1] for a in ("u", "v", "w")], axis=-1)
    J, K = slice(1, nj), slice(1, nk)
    J1, K1 = slice(2, nj + 1), slice(2, nk + 1)
    tris = ((P[J, K], P[J1, K], P[J, K1]), (P[J1, K1], P[J, K1], P[J, K]))
    area, parts, dns = 0.0, [], []
    for a, b, c in tris:
        n = np.cross(b - a, c - a)
        ar = 0.5 * np.linalg.norm(n, axis=-1)
        n = n / np.linalg.norm(n, axis=-1)[..., None]
        dn = np.abs(np.sum((C[J, K] - a) * n, axis=-1))
        ut = U[J, K] - np.sum(U[J, K] * n, axis=-1)[..., None] * n
        parts.append((np.sqrt(re * np.linalg.norm(ut, axis=-1) / dn) * dn, ar))
        dns.append((dn, ar))
        area = area + ar
    want = sum(y * ar for y, ar in parts) / area
    got = o.array("yplus")[1, J, K]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # a slip velocity along the span (the cylinder's axis) gives sqrt(Re |w| d_n) per triangle
    z = np.zeros(o.shape, order="F")
    o.set_state(z, z, np.full(o.shape, 0.7, order="F"), z, z, z, z)
    o.calccoef()
    want = sum(np.sqrt(re * 0.7 * dn) * ar for dn, ar in dns) / area
    got = o.array("yplus")[1, J, K]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_lambda2_closed_forms():
    """lambda_2 and |omega| of a linear velocity field u = A x: the discrete gradient is A to
    rounding on the interior, so lambda_2 is the middle eigenvalue of S^2 + Omega^2
    (S, Omega the symmetric and antisymmetric parts of A) and |omega| the norm of the curl."""
    g = load_golden("rles_9x21x5_les_converged")
    o = make(g)
    x, y, zc = o.array("x"), o.array("y"), o.array("z")
    A = np.array([[0.3, 0.8, 0.1], [-0.5, -0.2, 0.4], [0.2, -0.7, -0.1]])
    vel = [np.asfortranarray(A[a, 0] * x + A[a, 1] * y + A[a, 2] * zc) for a in range(3)]
    z = np.zeros(o.shape, order="F")
    o.set_state(vel[0], vel[1], vel[2], z, z, z, z)
    o.set_temperature(z, z)
    o.writedata_fields()
    S, Om = 0.5 * (A + A.T), 0.5 * (A - A.T)
    lam = np.sort(np.linalg.eigvalsh(S @ S + Om @ Om))
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    ni, nj, nk = int(g["ni"]), int(g["nj"]), int(g["nk"])
    I = (slice(2, ni - 1), slice(1, nj), slice(1, nk))
    assert np.abs(o.array("lamda2")[I] - lam[1]).max() < 1e-9
    assert np.abs(o.array("vorm")[I] - np.linalg.norm(curl)).max() < 1e-9


POST_SUMS = ("u", "v", "w", "p", "t", "uu", "uv", "uw", "vv", "vw", "ww", "sgs11", "sgs12",
             "sgs13", "sgs22", "sgs23", "sgs33")


def run_post(g):
    """Step 1 at DTIME, then POST_PROCESS's averaging phase at DTIME_POST for npost steps."""
    o = make(g)
    o.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    steps, npost = int(g["scalars"].shape[0]), int(g["npost"])
    for s in range(1, steps + 1):
        if s == steps - npost + 1:
            o.post_begin()
            o.set_dtime(float(g["dt_post"]))
        o.step(int(g["nswpp"]), float(g["resimax"]), True)
    return o


def test_post_process_bitexact():
    """POST_PROCESS (post_process.f90:9-276) with the accumulation of CALCCOEF
    (calccoef.f90:286-307): the running sums after three averaging steps, the in-place means
    and Reynolds stresses, the node averages and the wall Cf / Cp / Nu equal the reference's
    arrays bit for bit, and its Tecplot files to their 9 printed digits."""
    g = load_golden("rles_9x21x5_les_post")
    o = run_post(g)
    for t in ("u", "v", "w", "pref", "t", "nut"):
        assert np.array_equal(o.array("p" if t == "pref" else t), g["s4_" + t]), t
    assert o.scalar("niter_pos") == int(g["niter_pos"]) == 3
    for q in POST_SUMS:
        assert np.array_equal(o.array("sum_" + q), g["post_sum_" + q]), q
    o.post_finalize()
    for q in ("u", "v", "w", "p", "t", "uu", "uv", "uw", "vv", "vw", "ww"):
        assert np.array_equal(o.array("sum_" + q), g["post_fin_" + q]), q
    n = tuple(slice(1, m) for m in o.shape)
    for q, name in (("pg", "avg_pg"), ("tg", "avg_tg"), ("uu", "rs_uu"), ("uv", "rs_uv"),
                    ("uw", "rs_uw"), ("vv", "rs_vv"), ("vw", "rs_vw"), ("ww", "rs_ww")):
        assert np.array_equal(o.array(name)[n], g["post_node_" + q]), q
    w = (1, slice(1, o.nj), slice(1, o.nk))
    for q in ("cf", "cp", "nu"):
        assert np.array_equal(o.array(q + "_s")[w], g["post_surf_" + q]), q
    for q, name in (("pg", "avg_pg"), ("ug", "avg_ug"), ("vg", "avg_vg"), ("wg", "avg_wg"),
                    ("tg", "avg_tg"), ("urms", "rms_ug"), ("vrms", "rms_vg"),
                    ("wrms", "rms_wg"), ("uu", "rs_uu"), ("uv", "rs_uv"), ("uw", "rs_uw"),
                    ("vv", "rs_vv"), ("vw", "rs_vw"), ("ww", "rs_ww")):
        np.testing.assert_allclose(o.array(name)[n], g["plt_" + q], rtol=1e-8, atol=1e-300,
                                   err_msg=q)
    for q in ("cf", "cp", "nu"):
        np.testing.assert_allclose(o.array(q + "_s")[w], g["plt_surf_" + q], rtol=1e-8,
                                   err_msg=q)


def test_post_process_closed_forms():
    """A state held fixed over the averaging window: the means are the state, the rms is |u|,
    the resolved Reynolds stresses vanish (only the SGS part is left), and on the wall Cp is
    twice the 8-cell pressure average.  Checked on a seeded state by calling the finalization
    after N identical accumulations (three CALCCOEF passes on the same fields)."""
    g = load_golden("rles_9x21x5_les_post")
    o = make(g)
    o.set_temperature(g["in_t"], g["in_adv"], float(g["pe"]))
    o.post_begin()
    for _ in range(3):
        o.calccoef()
    u, v, w = (o.array(c).copy() for c in ("u", "v", "w"))
    o.post_finalize()
    np.testing.assert_allclose(o.array("sum_u"), u, rtol=1e-15, atol=0)
    # nu_t = 0 before any SGS pass: the stresses are <u u> - <u><u> = 0 up to round-off
    for q in ("uu", "uv", "uw", "vv", "vw", "ww"):
        assert np.abs(o.array("sum_" + q)).max() < 1e-14 * np.abs(u).max() ** 2, q
    n = tuple(slice(1, m) for m in o.shape)
    P = o.array("sum_p")
    uabs = np.sqrt(u * u)
    ref = 0.125 * (uabs[1:, 1:, 1:] + uabs[1:, :-1, 1:] + uabs[:-1, 1:, 1:] + uabs[:-1, :-1, 1:]
                   + uabs[1:, 1:, :-1] + uabs[1:, :-1, :-1] + uabs[:-1, 1:, :-1]
                   + uabs[:-1, :-1, :-1])
    np.testing.assert_allclose(o.array("rms_ug")[n], ref, rtol=1e-14)
    cp = 0.25 * (P[1, 1:-1, 1:-1] + P[1, :-2, 1:-1] + P[0, 1:-1, 1:-1] + P[0, :-2, 1:-1]
                 + P[1, 1:-1, 2:] + P[1, :-2, 2:] + P[0, 1:-1, 2:] + P[0, :-2, 2:])
    np.testing.assert_allclose(o.array("cp_s")[1, 1:-1, 1:-1], cp, rtol=1e-13, atol=1e-300)
