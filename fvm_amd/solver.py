"""Host-side mirror of the reference time loop over the C ABI (RLES/main.f90:35-161).

Solver3D owns one device context; its methods keep the reference's subroutine names so a
driver reads like the reference's MAIN.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import capi


class Solver3D:
    """3D fractional 4-step solver on one GPU (device-resident state)."""

    def __init__(self, ni, nj, nk, re, dtime, xg, yg, zg, ref_a=math.pi, use_les=True,
                 nswpp=20, resimax=1e-4, poisson_mode=0, check_every=1, af_mode=0,
                 af_tol=1e-14):
        self.L = capi.load()
        self.ni, self.nj, self.nk = ni, nj, nk
        self.shape = (ni + 1, nj + 1, nk + 1)
        self.prm = capi.Params(ni, nj, nk, float(re), float(dtime), float(ref_a),
                               1 if use_les else 0, int(nswpp), float(resimax),
                               int(poisson_mode), int(check_every), int(af_mode),
                               float(af_tol))
        arrs = [np.asfortranarray(a, dtype=np.float64).reshape((ni, nj, nk), order="F")
                for a in (xg, yg, zg)]
        flat = [np.ascontiguousarray(a.ravel(order="F")) for a in arrs]
        h = ctypes.c_void_p()
        capi.check(self.L.fvm_create(ctypes.byref(self.prm), *[capi._dptr(a) for a in flat],
                                     ctypes.byref(h)), "fvm_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # restart I/O ---------------------------------------------------------------------
    def set(self, name, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        flat = np.ascontiguousarray(a.ravel(order="F"))
        capi.check(self.L.fvm_set_field(self.h, capi.FIELDS[name], capi._dptr(flat)), "set " + name)

    def get(self, name):
        flat = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        capi.check(self.L.fvm_get_field(self.h, capi.FIELDS[name], capi._dptr(flat)), "get " + name)
        return flat.reshape(self.shape, order="F")

    def set_state(self, u, v, w, p, nlx, nly, nlz):
        for n, a in zip(("u", "v", "w", "p", "nlx", "nly", "nlz"), (u, v, w, p, nlx, nly, nlz)):
            self.set(n, a)

    # the time step ----------------------------------------------------------------------
    def step(self):
        info = capi.StepInfo()
        code = self.L.fvm_step(self.h, ctypes.byref(info))
        if code not in (0, 4):
            raise capi.FvmError(code, "fvm_step")
        d = info.as_dict()
        d["status"] = code
        return d

    def sync(self):
        capi.check(self.L.fvm_sync(self.h), "fvm_sync")

    def set_dtime(self, dtime: float):
        """Change DTIME between steps (the reference reassigns it at run time)."""
        capi.check(self.L.fvm_set_dtime(self.h, float(dtime)), "fvm_set_dtime")
        self.dtime = float(dtime)

    # the subroutines of the step (reference names) ------------------------------------
    def begin_step(self): capi.check(self.L.fvm_begin_step(self.h), "begin_step")
    def contra_vel(self): capi.check(self.L.fvm_contra_vel(self.h), "contra_vel")
    def nl(self): capi.check(self.L.fvm_nl(self.h), "nl")
    def dif(self): capi.check(self.L.fvm_dif(self.h), "dif")
    def p_source(self): capi.check(self.L.fvm_p_source(self.h), "p_source")
    def iter_u(self): capi.check(self.L.fvm_iter_u(self.h), "iter_u")
    def correct(self, sign): capi.check(self.L.fvm_correct(self.h, float(sign)), "correct")

    def stress_tensor_dy(self):
        c = ctypes.c_double()
        capi.check(self.L.fvm_stress_tensor_dy(self.h, ctypes.byref(c)), "stress_tensor_dy")
        return c.value

    def iter_p(self):
        r, n = ctypes.c_double(), ctypes.c_int()
        code = self.L.fvm_iter_p(self.h, ctypes.byref(r), ctypes.byref(n))
        if code not in (0, 4):
            raise capi.FvmError(code, "iter_p")
        return r.value, n.value

    def pressure_reference(self):
        r = ctypes.c_double()
        capi.check(self.L.fvm_pressure_reference(self.h, ctypes.byref(r)), "pressure_reference")
        return r.value

    def calccoef(self):
        info = capi.StepInfo()
        capi.check(self.L.fvm_calccoef(self.h, ctypes.byref(info)), "calccoef")
        return info.as_dict()

    # kernel timing ----------------------------------------------------------------------
    def profile(self, on=True):
        capi.check(self.L.fvm_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_reset(self):
        capi.check(self.L.fvm_profile_reset(self.h), "profile_reset")

    def profile_get(self):
        cap = 64
        names = ctypes.create_string_buffer(48 * cap)
        ms = (ctypes.c_double * cap)()
        calls = (ctypes.c_long * cap)()
        n = self.L.fvm_profile_get(self.h, names, ms, calls, cap)
        out = {}
        for q in range(max(n, 0)):
            nm = names.raw[48 * q:48 * q + 48].split(b"\0", 1)[0].decode()
            out[nm] = (ms[q], calls[q])
        return out
