"""fvm_calccoef_noalloc (CALCCOEF with its partial sums in a static device array, no device
allocation) gives exactly fvm_calccoef's results, on two contexts in a row, the second one
smaller, so the second call reuses the array the first call left behind."""
import ctypes
import math

import numpy as np
import pytest

from fvm_amd import capi

pytestmark = pytest.mark.gpu

KEYS = ("cfl", "cpb", "cd", "cl", "cd_p", "cd_v", "cl_p", "cl_v", "c_avg")


def _stepped(ni, nj, nk):
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    (xg, yg, zg), ic = make_case(ni, nj, nk, h1=0.01, r_out=8.0, sigma=0.05)
    s = Solver3D(ni, nj, nk, 3900.0, 2e-4, xg, yg, zg, ref_a=math.pi, use_les=True, nswpp=50,
                 resimax=1e-10, poisson_mode=3)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    s.step()
    return s


def _coef(s, fn):
    info = capi.StepInfo()
    capi.check(getattr(s.L, fn)(s.h, ctypes.byref(info)), fn)
    return info.as_dict()


@pytest.mark.parametrize("dims", [[(65, 97, 33), (17, 33, 9)]])
def test_noalloc_calccoef_equals_calccoef(dims):
    for ni, nj, nk in dims:
        print(f"calccoef_noalloc on {ni}x{nj}x{nk} nodes", flush=True)
        s = _stepped(ni, nj, nk)
        a = _coef(s, "fvm_calccoef_noalloc")
        b = _coef(s, "fvm_calccoef")
        c = _coef(s, "fvm_calccoef_noalloc")
        for k in KEYS:
            assert a[k] == b[k] == c[k], (ni, k, a[k], b[k], c[k])
        assert np.isfinite(a["cd"]) and a["cfl"] > 0
        s.close()


def test_noalloc_calccoef_2d():
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver2D
    (xg, yg, _), ic = make_case(33, 65, 1, h1=0.01, r_out=8.0, sigma=0.05)
    s = Solver2D(33, 65, 100.0, 5e-3, xg[:, :, 0], yg[:, :, 0], nswpp=50, resimax=1e-10,
                 poisson_mode=3)
    s.set_state(ic["u"][:, :, 0], ic["v"][:, :, 0], ic["p"][:, :, 0], ic["nlx"][:, :, 0],
                ic["nly"][:, :, 0])
    s.step()
    a = _coef(s, "fvm_calccoef_noalloc")
    b = _coef(s, "fvm_calccoef")
    for k in KEYS[:8]:
        assert a[k] == b[k], (k, a[k], b[k])
