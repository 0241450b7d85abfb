"""The batched 2D solver's set-up (host code of the HIP library, no GPU needed) against the
flang-built 2D reference's own set-up arrays (golden fixture): cell centres, Jacobian, metrics
and angles of GRID / METRIC / INIT (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/
grid.f90:47-115, metric.f90:18-145, init.f90:71-99), bit for bit."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_golden

LIB = os.path.join(REPO, "fvm_amd", "libfvm_amd.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfvm_amd.so not built")
def test_geometry2d_matches_reference_setup():
    g = load_golden("r2d_17x33_re100_converged")
    ni, nj = int(g["ni"]), int(g["nj"])
    L = ctypes.CDLL(LIB)
    pd = ctypes.POINTER(ctypes.c_double)
    L.fvm_geometry2d.argtypes = [ctypes.c_int, ctypes.c_int, pd, pd, pd]
    xg = np.ascontiguousarray(np.asarray(g["xg"]).ravel(order="F"))
    yg = np.ascontiguousarray(np.asarray(g["yg"]).ravel(order="F"))
    n = (ni + 1) * (nj + 1)
    out = np.zeros(9 * n)
    assert L.fvm_geometry2d(ni, nj, xg.ctypes.data_as(pd), yg.ctypes.data_as(pd),
                            out.ctypes.data_as(pd)) == 0
    planes = out.reshape(9, nj + 1, ni + 1).transpose(0, 2, 1)  # [q, i, j]
    for q, t in enumerate(("x", "y", "djac", "xix", "xiy", "etx", "ety", "theta")):
        ref = g["setup_" + t]
        assert np.array_equal(planes[q], ref), (t, float(np.abs(planes[q] - ref).max()))


def test_calccoef2d_checker_matches_reference_scalars():
    """The numpy 2D CALCCOEF used as the force checker reproduces the flang reference's C_D,
    C_L, their parts and C_pb from the reference's own fields."""
    from calccoef2d_np import calccoef2d
    g = load_golden("r2d_17x33_re100_converged")
    for st in (1, 2, 3):
        f = calccoef2d(g["xg"], g["yg"], g["setup_x"], g["setup_y"], g["setup_theta"],
                       g[f"s{st}_u"], g[f"s{st}_v"], g[f"s{st}_p"], float(g["re"]))
        row = g["scalars"][st - 1]  # step niter resip cfl cpb cd cl cd_p cd_v cl_p cl_v
        for k, col in (("cpb", 4), ("cd", 5), ("cl", 6), ("cd_p", 7), ("cd_v", 8), ("cl_p", 9),
                       ("cl_v", 10)):
            assert abs(f[k] - row[col]) <= 1e-13 * max(abs(row[col]), 1.0), (st, k, f[k], row[col])
