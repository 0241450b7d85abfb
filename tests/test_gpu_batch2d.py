"""The batched native 2D solver (fvm_batch2d_*) on the GPU.

* Against the flang-built 2D reference (golden fixture, 16 x 32 cells, Re 100, three steps with
  the pressure converged): u, v, p within 1e-10 of the field maximum, NL and DIF of the first
  step, C_D, C_L, their parts, C_pb and CFL within 1e-9 -- with the FP64 and the mixed Poisson
  correction.
* A case's results do not depend on the other cases of its batch (bit for bit).
* Long circumferential lines (NJ = 385, 513: the J-line scans with six and eight rows per lane
  and the multi-chunk wall sums of CALCCOEF) against the CPU oracle on the 2D slab: the
  oracle's residual of ITER_P for the GPU's pressure, and the oracle's projection with it,
  reproduce the GPU's step; C_D and C_L against the numpy CALCCOEF.
* The 96-case ensemble stepped in one batch equals every case stepped alone, bit for bit.
* The autotuner leaves the state as it found it.
"""
import math

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

LZ = 1.0e6


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def golden_batch(mixed, extra=()):
    from fvm_amd.batch2d import Batch2D
    g = load_golden("r2d_17x33_re100_converged")
    spec = dict(ni=int(g["ni"]), nj=int(g["nj"]), re=float(g["re"]), dt=float(g["dt"]),
                xg=g["xg"], yg=g["yg"])
    b = Batch2D([spec] + list(extra), mixed=mixed, resimax=1e-14, npass=8)
    b.set_state(0, g["in_u"], g["in_v"], g["in_p"], g["in_nlx"], g["in_nly"])
    return b, g


@pytest.mark.parametrize("mixed", [False, True])
def test_batch2d_matches_2d_reference(mixed):
    b, g = golden_batch(mixed)
    I = (slice(1, int(g["ni"])), slice(1, int(g["nj"])))
    for st in (1, 2, 3):
        info = b.step(1)[0]
        print(f"batch2d mixed={mixed} step {st}: passes {info['niter']} resip "
              f"{info['resip']:.2e}", flush=True)
        assert info["status"] == 0 and info["resip"] <= 1e-14, info
        for c in ("u", "v", "p"):
            assert rel(b.get(0, c), g[f"s{st}_{c}"]) < 1e-10, (st, c)
        assert rel(b.get(0, "nlx")[I], g[f"s{st}_nlx"][I]) < 1e-12, st
        if st == 1:
            assert rel(b.get(0, "difx")[I], g["s1_difx"][I]) < 1e-13
        row = g["scalars"][st - 1]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        for key, col in (("cfl", 3), ("cpb", 4), ("cd", 5), ("cl", 6), ("cd_p", 7),
                         ("cd_v", 8), ("cl_p", 9), ("cl_v", 10)):
            ref = row[col]
            assert abs(info[key] - ref) <= 1e-9 * max(abs(ref), 1.0), (st, key, info[key], ref)


def test_case_independent_of_its_batch():
    from fvm_amd.grid import ensemble_cases
    others = ensemble_cases(3)
    extra = [dict(ni=c["ni"], nj=c["nj"], re=c["re"], dt=c["dt"], xg=c["xg"][:, :, 0],
                  yg=c["yg"][:, :, 0]) for c in others]
    alone, _ = golden_batch(False)
    batch, _ = golden_batch(False, extra)
    for q, c in enumerate(others):
        batch.set_state(q + 1, *(c["ic"][k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly")))
    alone.step(3)
    batch.step(3)
    for c in ("u", "v", "p", "nlx", "nly"):
        assert np.array_equal(alone.get(0, c), batch.get(0, c)), c


def _slab(a2, planes):
    return np.asfortranarray(np.repeat(np.asarray(a2)[:, :, None], planes, axis=2))


@pytest.mark.parametrize("ni,nj", [(65, 513), (129, 385)])
def test_long_lines_against_oracle_slab(ni, nj):
    from calccoef2d_np import calccoef2d
    from fvm_amd.batch2d import Batch2D
    from fvm_amd.grid import make_case
    from oracle.oracle3d import Oracle3D
    re, dt = 180.0, 2.5e-3
    (xg, yg, _), ic = make_case(ni, nj, 1, h1=0.002, r_out=30.0, sigma=1e-3, seed=77)
    xg, yg = xg[:, :, 0], yg[:, :, 0]
    st = [ic[k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly")]
    b = Batch2D([dict(ni=ni, nj=nj, re=re, dt=dt, xg=xg, yg=yg)], mixed=False, resimax=1e-12,
                npass=6)
    b.set_state(0, *st)
    print(f"long lines {ni}x{nj}", flush=True)
    info = b.step(1)[0]
    assert info["status"] == 0 and info["resip"] <= 1e-12, info
    zg = np.zeros((ni, nj, 2), order="F")
    zg[:, :, 1] = LZ
    o = Oracle3D(ni, nj, 2, re, dt, LZ, _slab(xg, 2), _slab(yg, 2), zg)
    o.set_no_pref(True)
    z = np.zeros((ni + 1, nj + 1, 3), order="F")
    o.set_state(_slab(st[0], 3), _slab(st[1], 3), z, _slab(st[2], 3), _slab(st[3], 3),
                _slab(st[4], 3), z)
    o.predictor(False)
    o.correct_predictor()
    o.array("p")[:, :, :] = _slab(b.get(0, "p"), 3)
    res = o.p_residual()
    print(f"GPU resip {info['resip']:.3e}, oracle residual of the GPU pressure {res:.3e}",
          flush=True)
    assert res <= 1e-10, res
    o.finish_step()
    for c in ("u", "v", "p"):
        assert rel(b.get(0, c), o.array(c)[:, :, 1]) < 1e-10, c
    f = calccoef2d(xg, yg, b.get(0, "x"), b.get(0, "y"), b.get(0, "theta"), b.get(0, "u"),
                   b.get(0, "v"), b.get(0, "p"), re)
    for k in ("cd", "cl", "cd_p", "cd_v", "cl_p", "cl_v", "cpb"):
        assert abs(info[k] - f[k]) <= 1e-12 * max(abs(f[k]), 1.0), (k, info[k], f[k])


def test_ensemble_96_cases_equal_single_case_runs():
    from fvm_amd.batch2d import batch_from_cases
    from fvm_amd.grid import ensemble_cases
    cases = ensemble_cases(96)
    print("ensemble of", len(cases), "cases", flush=True)
    ens = batch_from_cases(cases, mixed=True, resimax=1e-8, npass=3, dt_scale=0.1)
    infos = ens.step(2)
    assert all(np.isfinite(i["cd"]) for i in infos)
    for q, c in enumerate(cases):
        one = batch_from_cases([c], mixed=True, resimax=1e-8, npass=3, dt_scale=0.1)
        i1 = one.step(2)[0]
        assert i1["cd"] == infos[q]["cd"] and i1["resip"] == infos[q]["resip"], q
        for f in ("u", "v", "p"):
            assert np.array_equal(one.get(0, f), ens.get(q, f)), (q, f)
        one.close()


def test_autotune_restores_state_and_results():
    from fvm_amd.batch2d import batch_from_cases
    from fvm_amd.grid import ensemble_cases
    cases = ensemble_cases(6)
    a = batch_from_cases(cases, mixed=True, resimax=1e-8, npass=3, dt_scale=0.1)
    b = batch_from_cases(cases, mixed=True, resimax=1e-8, npass=3, dt_scale=0.1)
    a.step(1)
    b.step(1)
    before = [a.get(q, "u") for q in range(len(cases))]
    res = a.autotune(reps=2)
    print("autotune", res, flush=True)
    assert res["tile_j"] in (1, 2, 4, 8) and res["fft_cols"] in (1, 2, 4, 8, 16)
    for q in range(len(cases)):
        assert np.array_equal(a.get(q, "u"), before[q])
    a.step(2)
    b.step(2)
    for q in range(len(cases)):
        for f in ("u", "v", "p"):
            assert rel(a.get(q, f), b.get(q, f)) < 1e-12, (q, f)
    h = a.history(2)
    assert h.shape == (2, len(cases), 8) and np.all(np.isfinite(h))
