"""The ensemble step (fvm_ensemble_step: independent 2D cases of uneven size and Reynolds
number, one stream per case, several host threads) gives every case exactly the fields of its
own fvm_step sequence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ensemble_step_equals_single_case_steps():
    from fvm_amd.grid import ensemble_cases
    from fvm_amd.solver import ensemble2d
    cases = ensemble_cases(5)
    print("ensemble:", [(c["ni"], c["nj"], round(c["re"], 1)) for c in cases], flush=True)
    kw = dict(nswpp=50, resimax=1e-10, poisson_mode=3)
    ens = ensemble2d(cases, nthreads=3, **kw)
    alone = ensemble2d(cases, nthreads=1, **kw)
    for st in range(2):
        infos = ens.step()
        ref = [s.step() for s in alone.solvers]
        for q, (a, b) in enumerate(zip(infos, ref)):
            assert a["resip"] <= 1e-10 and np.isfinite(a["cd"]), (st, q, a)
            assert a["cd"] == b["cd"] and a["cl"] == b["cl"] and a["niter"] == b["niter"]
    for q, (s, t) in enumerate(zip(ens.solvers, alone.solvers)):
        for c in ("u", "v", "p"):
            assert np.array_equal(s.get(c), t.get(c)), (q, c)
