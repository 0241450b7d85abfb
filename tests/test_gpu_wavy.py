"""GPU parity on a grid that is not an extruded O-grid (golden fixture of the flang-built
reference): a wavy cylinder whose radius varies along the span, so xi_z, eta_z and the stage
cross couplings of the AF step do not vanish and the Poisson face coefficients vary along J
and K.  This runs the general code paths the benchmark grid never takes: the metric kernels
with all nine metrics, the chunked AF line iteration (the line scans and the independent K-line
solve need vanishing couplings), and the pressure defect correction on a grid where the
separable inverse is only approximate."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

NAME = "rles_13x17x9_wavy_converged"


def rel(a, b, sl=(slice(None),) * 3):
    a, b = a[sl], b[sl]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make_gpu(g, **kw):
    from fvm_amd.solver import Solver3D
    s = Solver3D(int(g["ni"]), int(g["nj"]), int(g["nk"]), float(g["re"]), float(g["dt"]),
                 g["xg"], g["yg"], g["zg"], ref_a=float(g["ref_a"]), use_les=True, **kw)
    s.set_state(*(g["in_" + k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    return s


def inner(g):
    return (slice(1, int(g["ni"])), slice(1, int(g["nj"])), slice(1, int(g["nk"])))


def test_wavy_setup_and_stages():
    g = load_golden(NAME)
    s = make_gpu(g, nswpp=50, resimax=1e-13, poisson_mode=3)
    for t in ("x", "y", "z", "djac", "xix", "xiy", "xiz", "etx", "ety", "etz", "ztx", "zty",
              "ztz"):
        assert rel(s.get(t), g["setup_" + t]) < 1e-15, t
    assert np.abs(g["setup_xiz"]).max() > 1e-3  # the grid is not extruded
    I = inner(g)
    s.begin_step()
    s.contra_vel()
    s.nl()
    assert rel(s.get("nlx"), g["s1_nl0x"], I) < 1e-14
    s.stress_tensor_dy()
    assert rel(s.get("nut"), g["s1_nut"]) < 1e-12
    for c in ("sgsx", "sgsy", "sgsz", "nlx", "nly", "nlz"):
        assert rel(s.get(c), g["s1_" + c], I) < 1e-12, c
    s.dif()
    for c in ("difx", "dify", "difz"):
        assert rel(s.get(c), g["s1_" + c], I) < 1e-14, c
    s.p_source()
    assert rel(s.get("s_u"), g["s1_su0"], I) < 1e-14
    s.iter_u()
    for c, t in (("u", "ustar"), ("v", "vstar"), ("w", "wstar")):
        assert rel(s.get(c), g["s1_" + t], I) < 1e-12, c
    s.correct(-1.0)
    assert rel(s.get("u"), g["s1_uhat"]) < 1e-12


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_wavy_converged_steps(mode):
    g = load_golden(NAME)
    s = make_gpu(g, nswpp=200000 if mode == 0 else 200, resimax=1e-14, poisson_mode=mode,
                 check_every=4)
    for st in (1, 2):
        info = s.step()
        print(f"wavy mode {mode} step {st}: passes {info['niter']} resip {info['resip']:.2e}",
              flush=True)
        assert info["status"] == 0 and info["resip"] <= 1e-14, info
        for c in ("u", "v", "w"):
            assert rel(s.get(c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        assert rel(s.get("p"), g[f"s{st}_pref"]) < 1e-10, (st, "p")
        row = g["scalars"][st - 1]  # step niter resip pref cfl cpb cd cl cd_p cd_v
        assert abs(info["cd"] - row[6]) <= 1e-9 * abs(row[6])
        assert abs(info["cl"] - row[7]) <= 1e-9 * max(abs(row[7]), 1.0)
