"""ctypes binding of the C ABI in include/fvm_capi.h.

The HIP library is the only compute path: loading fails loudly when fvm_amd/libfvm_amd.so is
missing or no GPU is visible; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvm_amd.so")

FIELDS = {
    "u": 0, "v": 1, "w": 2, "p": 3, "nlx": 4, "nly": 5, "nlz": 6, "difx": 7, "dify": 8,
    "difz": 9, "s_u": 10, "s_v": 11, "s_w": 12, "djac": 13, "xix": 14, "etx": 15, "ztx": 16,
    "xiy": 17, "ety": 18, "zty": 19, "xiz": 20, "etz": 21, "ztz": 22, "x": 23, "y": 24, "z": 25,
    "theta": 26, "nut": 27, "sgsx": 28, "sgsy": 29, "sgsz": 30, "un": 31,
    "vn": 32, "wn": 33, "t": 34, "adv": 35, "dift": 36, "sgst": 37, "kappa": 38, "that": 39,
    "dtdx": 40, "dtdy": 41, "dtdz": 42, "tn": 43, "vorm": 44, "lamda2": 45, "pg": 46,
    "ug": 47, "vg": 48, "wg": 49, "tg": 50, "vorg": 51, "lamdag": 52,
    "yplus": 53,
}
# time averages: the running sums (the means / Reynolds stresses after the finalization), the
# node averages of the Avrg, Urms and Strs files and the wall coefficients of the Surf file
POST_SUMS = ("u", "v", "w", "p", "t", "uu", "uv", "uw", "vv", "vw", "ww", "sgs11", "sgs12",
             "sgs13", "sgs22", "sgs23", "sgs33")
POST_OUT = ("avg_pg", "avg_ug", "avg_vg", "avg_wg", "avg_tg", "rms_ug", "rms_vg", "rms_wg",
            "rs_uu", "rs_uv", "rs_uw", "rs_vv", "rs_vw", "rs_ww", "surf_cf", "surf_cp",
            "surf_nu")
FIELDS.update({"sum_" + q: 54 + i for i, q in enumerate(POST_SUMS)})
FIELDS.update({q: 71 + i for i, q in enumerate(POST_OUT)})

STATUS = {0: "ok", 1: "bad argument", 2: "HIP runtime error", 3: "non-finite value",
          4: "Poisson sweep cap reached", 5: "grid too large for a line kernel"}


class FvmError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: status {code} ({STATUS.get(code, '?')})")


class Params(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int),
                ("re", ctypes.c_double), ("dtime", ctypes.c_double), ("ref_a", ctypes.c_double),
                ("use_les", ctypes.c_int), ("nswpp", ctypes.c_int), ("resimax", ctypes.c_double),
                ("poisson_mode", ctypes.c_int), ("check_every", ctypes.c_int),
                ("af_mode", ctypes.c_int), ("af_tol", ctypes.c_double)]


class StepInfo(ctypes.Structure):
    _fields_ = [("resip", ctypes.c_double), ("niter", ctypes.c_int), ("nouter", ctypes.c_int),
                ("pref", ctypes.c_double), ("c_avg", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cpb", ctypes.c_double), ("cd", ctypes.c_double), ("cl", ctypes.c_double),
                ("cd_p", ctypes.c_double), ("cd_v", ctypes.c_double), ("cl_p", ctypes.c_double),
                ("cl_v", ctypes.c_double), ("af_iters", ctypes.c_int * 3),
                ("nu", ctypes.c_double), ("c_t", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["af_iters"] = list(self.af_iters)
        return d


EXPORTS = ("fvm_create", "fvm_destroy", "fvm_set_field", "fvm_get_field", "fvm_step",
           "fvm_begin_step", "fvm_contra_vel", "fvm_nl", "fvm_stress_tensor_dy", "fvm_dif",
           "fvm_p_source", "fvm_iter_u", "fvm_correct", "fvm_iter_p", "fvm_pressure_reference",
           "fvm_calccoef", "fvm_divergence", "fvm_profile_enable", "fvm_profile_get",
           "fvm_profile_reset",
           "fvm_sync", "fvm_set_dtime", "fvm_create2d", "fvm_set_thermal", "fvm_iter_t",
           "fvm_writedata_fields", "fvm_post_begin", "fvm_post_finalize", "fvm_ensemble_step",
           "fvm_calccoef_noalloc", "fvm_copy_field", "fvm_nl_c", "fvm_dif_c",
           "fvm_stress_tensor_dy_c", "fvm_add_sgs_c", "fvm_iter_u_c", "fvm_correct_c", "fvm_bcu",
           "fvm_bcp", "fvm_batch2d_create", "fvm_batch2d_destroy",
           "fvm_batch2d_set_field", "fvm_batch2d_get_field", "fvm_batch2d_set_dtime",
           "fvm_batch2d_set_wall_velocity", "fvm_batch2d_step", "fvm_batch2d_history",
           "fvm_batch2d_set_tuning", "fvm_batch2d_get_tuning", "fvm_batch2d_autotune",
           "fvm_batch2d_profile",
           "fvm_batch2d_adjoint_init", "fvm_batch2d_adjoint_step", "fvm_batch2d_adjoint_wall",
           "fvm_batch2d_adjoint_get", "fvm_batch2d_save", "fvm_batch2d_load", "fvm_geometry2d")

_LIB = None


def load(path: str = LIB_PATH):
    """Load the HIP library and declare the C ABI signatures (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() "
                                "(make -C fvm_amd/csrc)")
    L = ctypes.CDLL(path)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    L.fvm_create.argtypes = [ctypes.POINTER(Params), pd, pd, pd, ctypes.POINTER(vp)]
    L.fvm_create2d.argtypes = [ctypes.POINTER(Params), pd, pd, ctypes.POINTER(vp)]
    L.fvm_destroy.argtypes = [vp]
    L.fvm_set_field.argtypes = [vp, ci, pd]
    L.fvm_get_field.argtypes = [vp, ci, pd]
    L.fvm_step.argtypes = [vp, ctypes.POINTER(StepInfo)]
    for f in ("fvm_begin_step", "fvm_contra_vel", "fvm_nl", "fvm_dif", "fvm_p_source",
              "fvm_iter_u", "fvm_sync", "fvm_profile_reset"):
        getattr(L, f).argtypes = [vp]
    L.fvm_stress_tensor_dy.argtypes = [vp, pd]
    L.fvm_correct.argtypes = [vp, cd]
    L.fvm_set_dtime.argtypes = [vp, cd]
    L.fvm_set_thermal.argtypes = [vp, cd]
    L.fvm_iter_t.argtypes = [vp]
    L.fvm_writedata_fields.argtypes = [vp]
    L.fvm_post_begin.argtypes = [vp, vp]
    L.fvm_post_finalize.argtypes = [vp, ctypes.POINTER(ci)]
    L.fvm_ensemble_step.argtypes = [ctypes.POINTER(vp), ci, ci, ctypes.POINTER(StepInfo)]
    L.fvm_iter_p.argtypes = [vp, pd, ctypes.POINTER(ci)]
    L.fvm_pressure_reference.argtypes = [vp, pd]
    L.fvm_calccoef.argtypes = [vp, ctypes.POINTER(StepInfo)]
    L.fvm_divergence.argtypes = [vp, pd]
    L.fvm_calccoef_noalloc.argtypes = [vp, ctypes.POINTER(StepInfo)]
    for f in ("fvm_copy_field", "fvm_nl_c", "fvm_dif_c", "fvm_add_sgs_c", "fvm_bcu"):
        getattr(L, f).argtypes = [vp, ci, ci]
    L.fvm_stress_tensor_dy_c.argtypes = [vp, pd]
    L.fvm_iter_u_c.argtypes = [vp, ci, ci, ci, ci, ci, ci]
    L.fvm_correct_c.argtypes = [vp, ci, ci, cd]
    L.fvm_bcp.argtypes = [vp]
    L.fvm_profile_enable.argtypes = [vp, ci]
    L.fvm_profile_get.argtypes = [vp, ctypes.c_char_p, pd, ctypes.POINTER(ctypes.c_long), ci]
    for f in EXPORTS:
        getattr(L, f).restype = ci
    _LIB = L
    return L


def check(code, where):
    if code != 0:
        raise FvmError(code, where)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
