"""The FP32 change form of the batched 2D AF sweeps (Batch2D af32) on the GPU.

Each stage solves T (x - x0) = r in FP32, r the FP64 residual of the stage input x0 against the
stage's line operator, and adds the change to the FP64 field.  The solution of the line system
is the one the FP64 sweeps find, so the form is checked against those:

* the flang-built 2D reference (golden fixture, 16 x 32 cells, three steps, pressure converged):
  u, v within 2e-7 and p within 2e-6 of their maximum (the FP64 sweeps: 1e-10);
* first step from a seeded start on cases whose lines span one to eight rows per lane: the
  difference to the FP64 sweeps within 2e-6 of the step's change of the field;
* 40 steps of a mixed batch: fields within 1e-6 of the FP64 sweeps', C_D within 1e-6;
* a case's results do not depend on the other cases of its batch (bit for bit);
* the adjoint refuses a change-form batch.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def specs_of(cases):
    return [dict(ni=c["ni"], nj=c["nj"], re=c["re"], dt=c["dt"], xg=c["xg"][:, :, 0],
                 yg=c["yg"][:, :, 0]) for c in cases]


def seeded(ni, nj, re, dt, seed):
    from fvm_amd.grid import make_case
    (xg, yg, _), ic = make_case(ni, nj, 1, h1=0.002, r_out=30.0, sigma=1e-3, seed=seed)
    return dict(ni=ni, nj=nj, re=re, dt=dt, xg=xg, yg=yg, ic=ic)


def batch(cases, af32, mixed=False, resimax=1e-12, npass=8):
    from fvm_amd.batch2d import Batch2D
    b = Batch2D(specs_of(cases), mixed=mixed, resimax=resimax, npass=npass, af32=af32)
    for q, c in enumerate(cases):
        b.set_state(q, *(c["ic"][k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly")))
    return b


def test_af32_matches_2d_reference():
    from fvm_amd.batch2d import Batch2D
    g = load_golden("r2d_17x33_re100_converged")
    spec = dict(ni=int(g["ni"]), nj=int(g["nj"]), re=float(g["re"]), dt=float(g["dt"]),
                xg=g["xg"], yg=g["yg"])
    b = Batch2D([spec], mixed=False, resimax=1e-14, npass=8, af32=True)
    b.set_state(0, g["in_u"], g["in_v"], g["in_p"], g["in_nlx"], g["in_nly"])
    worst = {}
    for st in (1, 2, 3):
        info = b.step(1)[0]
        assert info["status"] == 0 and info["resip"] <= 1e-14, info
        for c in ("u", "v", "p"):
            e = rel(b.get(0, c), g[f"s{st}_{c}"])
            print(f"step {st} {c}: FP32 change form vs the reference, rel {e:.2e}", flush=True)
            worst[c] = max(worst.get(c, 0.0), e)
    # the first step changes u by ~0.4 of its maximum: FP32 rounding of that change is ~2e-8
    assert worst["u"] < 2e-7 and worst["v"] < 2e-7 and worst["p"] < 2e-6, worst


@pytest.mark.parametrize("ni,nj", [(65, 129), (97, 257), (129, 385), (65, 513)])
def test_af32_first_step_against_fp64_sweeps(ni, nj):
    c = seeded(ni, nj, 150.0, 1e-3, seed=ni + nj)
    a, b = batch([c], af32=False), batch([c], af32=True)
    ia, ib = a.step(1)[0], b.step(1)[0]
    assert ia["status"] == 0 and ib["status"] == 0, (ia, ib)
    for f in ("u", "v"):
        x0 = c["ic"][f][:, :, 0]
        x64, x32 = a.get(0, f), b.get(0, f)
        ch = np.abs(x64 - x0)[1:ni, 1:nj].max()
        e = float(np.abs(x32 - x64)[1:ni, 1:nj].max() / ch)
        print(f"{ni}x{nj} {f}: |FP32 change form - FP64 sweeps| / |change| {e:.2e}", flush=True)
        assert e < 2e-6, f
    e = rel(b.get(0, "p"), a.get(0, "p"))
    print(f"{ni}x{nj} p: rel {e:.2e}", flush=True)
    assert e < 1e-5


def test_af32_steps_against_fp64_sweeps():
    from fvm_amd.grid import ensemble_cases
    cases = ensemble_cases(6, dt=1e-3, dt_rule="uniform")
    a = batch(cases, af32=False, mixed=True, resimax=1e-8, npass=3)
    b = batch(cases, af32=True, mixed=True, resimax=1e-8, npass=3)
    ia, ib = a.step(40), b.step(40)
    for q, c in enumerate(cases):
        assert ib[q]["status"] == 0, (q, ib[q])
        for f in ("u", "v", "p"):
            e = rel(b.get(q, f), a.get(q, f))
            print(f"case {q} ({c['ni']}x{c['nj']}) {f}: 40 steps, rel {e:.2e}", flush=True)
            assert e < 1e-6, (q, f)
        assert abs(ib[q]["cd"] - ia[q]["cd"]) <= 1e-6 * abs(ia[q]["cd"]), q


def test_af32_case_independent_of_its_batch():
    from fvm_amd.grid import ensemble_cases
    cases = ensemble_cases(4, dt=1e-3, dt_rule="uniform")
    ens = batch(cases, af32=True, mixed=True, resimax=1e-8, npass=3)
    ens.step(3)
    for q, c in enumerate(cases):
        one = batch([c], af32=True, mixed=True, resimax=1e-8, npass=3)
        one.step(3)
        for f in ("u", "v", "p"):
            assert np.array_equal(one.get(0, f), ens.get(q, f)), (q, f)
        one.close()


def test_af32_batch_refuses_adjoint():
    b = batch([seeded(33, 65, 100.0, 1e-3, seed=5)], af32=True)
    assert b.L.fvm_batch2d_adjoint_init(b.h) != 0
