"""The 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/) is the 3D step on a
one-cell span slab: the oracle run on the slab reproduces the flang-built 2D reference's
converged steps (golden fixture) to rounding.  Also the 2D post-processing (Strouhal number,
C_L amplitude, means and rms of calccoef.f90:156-195) on a synthetic lift signal."""
import math

import numpy as np

from conftest import load_golden
from oracle.oracle3d import Oracle3D

LZ = 1.0e6


def slab(a2, planes):
    return np.asfortranarray(np.repeat(np.asarray(a2)[:, :, None], planes, axis=2))


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_slab_oracle_matches_2d_reference():
    g = load_golden("r2d_17x33_re100_converged")
    ni, nj = int(g["ni"]), int(g["nj"])
    xg, yg = slab(g["xg"], 2), slab(g["yg"], 2)
    zg = np.zeros_like(xg)
    zg[:, :, 1] = LZ
    o = Oracle3D(ni, nj, 2, float(g["re"]), float(g["dt"]), LZ, xg, yg, np.asfortranarray(zg))
    o.set_no_pref(True)
    z = np.zeros((ni + 1, nj + 1, 3), order="F")
    o.set_state(slab(g["in_u"], 3), slab(g["in_v"], 3), z, slab(g["in_p"], 3),
                slab(g["in_nlx"], 3), slab(g["in_nly"], 3), z)
    I = (slice(1, ni), slice(1, nj))
    for t in ("xix", "xiy", "etx", "ety"):
        assert rel(o.array(t)[:, :, 1][I], g["setup_" + t][I]) < 1e-13, t
    for st in (1, 2, 3):
        o.step(200000, 1e-15, False)
        row = g["scalars"][st - 1]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        # same sweep sequence: the count differs by a few where the residual approaches 1e-15
        # at rounding level
        assert abs(o.scalar("niter") - row[1]) <= 10
        assert abs(o.scalar("cfl") - row[3]) <= 1e-12 * row[3]
        for c in ("u", "v", "p"):
            assert rel(o.array(c)[:, :, 1], g[f"s{st}_{c}"]) < 1e-12, (st, c)


def test_post_processing_strouhal_on_sinusoid():
    from fvm_amd.solver import Post2D
    f, amp, dt = 0.2, 0.3, 0.005
    post = Post2D()
    out = None
    for n in range(1, 4001):
        t = n * dt
        out = post.update(t, 1.3 + 0.02 * math.cos(4 * math.pi * f * t), amp * math.sin(2 * math.pi * f * t),
                          -0.9)
    # St = 1 / (time between extrema two apart) = f; amplitude = half the peak-to-peak
    assert abs(out["st"] - f) < 2 * dt * f * f / 0.5 + 1e-12
    assert abs(out["cl_amp"] - amp) < 1e-3
    assert abs(out["cd_mean"] - 1.3) < 1e-3
    assert abs(out["cl_rms"] - amp / math.sqrt(2.0)) < 2e-3
