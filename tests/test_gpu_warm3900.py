"""GPU against the flang-built reference at the benchmark's physics: Re 3900, wall spacing
0.001 D, outer radius 25 D, dt = 0.002 on a 32 x 64 x 16-cell grid.  The reference ran the
benchmark's start-up (20 steps at dt/10 from the seeded start), then two steps at dt, every
pressure solve to RESIP 1e-14 (golden fixture).  The GPU starts from the reference's state
after the start-up and takes the two steps at dt with each Poisson form."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

NAME = "rles_33x65x17_re3900_warm"


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("mode", [2, 3])
def test_two_steps_at_dt_after_startup(mode):
    from fvm_amd.solver import Solver3D
    g = load_golden(NAME)
    s = Solver3D(int(g["ni"]), int(g["nj"]), int(g["nk"]), float(g["re"]), float(g["dt"]),
                 g["xg"], g["yg"], g["zg"], ref_a=float(g["ref_a"]), use_les=True, nswpp=50,
                 resimax=1e-12, poisson_mode=mode)
    s.set_state(g["s20_u"], g["s20_v"], g["s20_w"], g["s20_pref"], g["s20_nlx"], g["s20_nly"],
                g["s20_nlz"])
    for st in (21, 22):
        info = s.step()
        print(f"mode {mode} step {st}: passes {info['niter']} resip {info['resip']:.2e} "
              f"cfl {info['cfl']:.3f}", flush=True)
        assert info["status"] == 0 and info["resip"] <= 1e-12, info
        for c in ("u", "v", "w"):
            if f"s{st}_{c}" in g:
                assert rel(s.get(c), g[f"s{st}_{c}"]) < TOL, (st, c)
        assert rel(s.get("p"), g[f"s{st}_pref"]) < TOL, (st, "p")
        row = g["scalars"][st - 1]  # step niter resip pref cfl cpb cd cl cd_p cd_v
        assert abs(info["cfl"] - row[4]) <= TOL * row[4]
        assert abs(info["cd"] - row[6]) <= TOL * abs(row[6])
        assert abs(info["cl"] - row[7]) <= TOL * max(abs(row[7]), 1.0)


TOL = 1e-9
