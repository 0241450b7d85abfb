"""ctypes wrapper of the C restatement of the reference 3D solver (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module; the
product package fvm_amd never imports it.  Arrays are (NI+1, NJ+1, NK+1) float64 in Fortran
order, ghosts included, exactly as the reference holds them with one MPI rank.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle3d.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, ci, cd, cs = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        L.orc3d_create.restype = vp
        L.orc3d_create.argtypes = [ci, ci, ci, cd, cd, cd]
        L.orc3d_destroy.argtypes = [vp]
        L.orc3d_array.restype = ctypes.POINTER(ctypes.c_double)
        L.orc3d_array.argtypes = [vp, cs]
        L.orc3d_scalar.restype = cd
        L.orc3d_scalar.argtypes = [vp, cs]
        L.orc3d_setup.argtypes = [vp, vp, vp, vp]
        L.orc3d_set_state.argtypes = [vp] * 8
        L.orc3d_step.argtypes = [vp, ci, cd, ci]
        L.orc3d_predictor.argtypes = [vp, ci]
        L.orc3d_post_begin.argtypes = [vp]
        L.orc3d_post_finalize.argtypes = [vp]
        L.orc3d_set_dtime.argtypes = [vp, cd]
        L.orc3d_set_no_pref.argtypes = [vp, ci]
        for f in ("orc3d_contra_vel", "orc3d_sgs", "orc3d_bcp", "orc3d_lisolv", "orc3d_calccoef",
                  "orc3d_pred_nl", "orc3d_pred_sgs", "orc3d_pred_dif_ps", "orc3d_pred_iter_u",
                  "orc3d_correct_predictor", "orc3d_finish_step", "orc3d_p_residual"):
            getattr(L, f).argtypes = [vp]
        L.orc3d_iter_p.argtypes = [vp, ci, cd]
        L.orc3d_nl.argtypes = [vp, vp, vp]
        L.orc3d_dif.argtypes = [vp, vp, vp]
        L.orc3d_dif_diagonal.argtypes = [vp, vp, vp]
        L.orc3d_p_source.argtypes = [vp, vp]
        L.orc3d_test_filter.argtypes = [vp, vp, vp]
        L.orc3d_bcu.argtypes = [vp, vp, vp]
        L.orc3d_mpi_x.argtypes = [vp, vp]
        L.orc3d_iter_u.argtypes = [vp] * 7
        L.orc3d_set_thermal.argtypes = [vp, ci]
        L.orc3d_set_pe.argtypes = [vp, cd]
        L.orc3d_set_temperature.argtypes = [vp, vp, vp]
        L.orc3d_iter_t.argtypes = [vp]
        L.orc3d_bct.argtypes = [vp, vp, vp]
        L.orc3d_writedata_fields.argtypes = [vp]
        _LIB = L
    return _LIB


def _full(nodes, ni, nj, nk):
    """Node array (NI, NJ, NK) -> oracle node layout (NI+1, NJ+1, NK+1) holding nodes 1..N."""
    a = np.zeros((ni + 1, nj + 1, nk + 1), order="F")
    a[1:, 1:, 1:] = nodes
    return a


class Oracle3D:
    """The reference 3D time step on the CPU, one rank, reference loop order."""

    def __init__(self, ni, nj, nk, re, dtime, ref_a, xg, yg, zg):
        self.L = lib()
        self.ni, self.nj, self.nk = ni, nj, nk
        self.shape = (ni + 1, nj + 1, nk + 1)
        self.h = ctypes.c_void_p(self.L.orc3d_create(ni, nj, nk, re, dtime, ref_a))
        nodes = [_full(np.asarray(a, dtype=np.float64), ni, nj, nk) for a in (xg, yg, zg)]
        self.L.orc3d_setup(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in nodes])

    def __del__(self):
        try:
            self.L.orc3d_destroy(self.h)
        except Exception:
            pass

    def array(self, name: str) -> np.ndarray:
        """Live view of a named oracle array (Fortran order, ghosts included)."""
        p = self.L.orc3d_array(self.h, name.encode())
        if not p:
            raise KeyError(name)
        n = int(np.prod(self.shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.shape, order="F")

    def scalar(self, name: str) -> float:
        return float(self.L.orc3d_scalar(self.h, name.encode()))

    def set_state(self, u, v, w, p, nlx, nly, nlz):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (u, v, w, p, nlx, nly, nlz)]
        self.L.orc3d_set_state(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])

    def set_temperature(self, t, adv, pe=2769.0):
        """Passive temperature T and its convective history ADV; enables ITER_T/BCT/SGS_T."""
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (t, adv)]
        self.L.orc3d_set_temperature(self.h, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs])
        self.L.orc3d_set_pe(self.h, float(pe))
        self.L.orc3d_set_thermal(self.h, 1)

    def writedata_fields(self):
        """|omega|, lambda_2 and the nodal averages of WRITEDATA for the current state."""
        self.L.orc3d_writedata_fields(self.h)

    def calccoef(self):
        """CALCCOEF on the current state: force coefficients, CFL, NU and the wall y+."""
        self.L.orc3d_calccoef(self.h)

    def set_no_pref(self, on=True):
        self.L.orc3d_set_no_pref(self.h, 1 if on else 0)

    def predictor(self, use_les=True):
        """The step up to and including the three ITER_U solves (u*, v*, w* left in u, v, w)."""
        self.L.orc3d_predictor(self.h, 1 if use_les else 0)

    # the predictor in its stages (RLES/main.f90:42-58, 59-69, 72-78, 82-88)
    def pred_nl(self):
        self.L.orc3d_pred_nl(self.h)

    def pred_sgs(self):
        self.L.orc3d_pred_sgs(self.h)

    def pred_dif_ps(self):
        self.L.orc3d_pred_dif_ps(self.h)

    def pred_iter_u(self):
        self.L.orc3d_pred_iter_u(self.h)

    def correct_predictor(self):
        """u^ = u* - dt S / J, BCU, CONTRA_VEL (RLES/main.f90:90-101)."""
        self.L.orc3d_correct_predictor(self.h)

    def p_residual(self):
        """ITER_P's scaled-L1 residual of the current p, no sweep (iter_p.f90:12-59)."""
        self.L.orc3d_p_residual(self.h)
        return self.scalar("resip")

    def finish_step(self):
        """BCP, P_SOURCE, projection, pressure reference, BCU, CALCCOEF (main.f90:105-150)."""
        self.L.orc3d_finish_step(self.h)

    def post_begin(self):
        """POST_PROCESS's start: zero the time-average sums; CALCCOEF then accumulates."""
        self.L.orc3d_post_begin(self.h)

    def post_finalize(self):
        """POST_PROCESS's end: means, rms, Reynolds stresses, node averages, Cf / Cp / Nu."""
        self.L.orc3d_post_finalize(self.h)

    def set_dtime(self, dtime):
        self.L.orc3d_set_dtime(self.h, float(dtime))

    def step(self, nswpp=20, resimax=1e-4, use_les=True):
        self.L.orc3d_step(self.h, int(nswpp), float(resimax), 1 if use_les else 0)
