"""Host-side mirror of the reference time loops over the C ABI.

Solver3D (RLES/main.f90:35-161) and Solver2D (cylinder/cylinder_v3.0.3_AF_original
4step_face_nl/Main.f90:48-110) own one device context each; their methods keep the reference's
subroutine names so a driver reads like the reference's MAIN.  Post2D is the 2D post-processing
phase of CALCCOEF (Strouhal number from C_L extrema, running means and rms).
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

    def set_temperature(self, t, adv, pe=2769.0):
        """Switch on the passive temperature (ITER_T, BCT, SGS_T, Nusselt number) with Peclet
        number pe and start it from T and its convective term ADV."""
        capi.check(self.L.fvm_set_thermal(self.h, float(pe)), "fvm_set_thermal")
        self.set("t", t)
        self.set("adv", adv)

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
    def iter_t(self): capi.check(self.L.fvm_iter_t(self.h), "iter_t")
    def writedata_fields(self):
        capi.check(self.L.fvm_writedata_fields(self.h), "writedata_fields")
    def correct(self, sign): capi.check(self.L.fvm_correct(self.h, float(sign)), "correct")

    def post_begin(self, t=None):
        """Start the time averages of POST_PROCESS (t: the temperature averaged without the
        passive temperature model)."""
        tp = None
        if t is not None:
            tp = np.ascontiguousarray(np.asfortranarray(t, dtype=np.float64).ravel(order="F"))
        capi.check(self.L.fvm_post_begin(self.h, tp.ctypes.data if tp is not None else None),
                   "post_begin")

    def post_finalize(self):
        """Means, rms, Reynolds stresses, node averages, wall coefficients; returns NITER_POS."""
        n = ctypes.c_int()
        capi.check(self.L.fvm_post_finalize(self.h, ctypes.byref(n)), "post_finalize")
        return n.value

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

    def divergence(self):
        """Face-flux divergence of the current velocity: max |div|, mean |div| and
        sum |div| / sum(|UE| + |VN| + |WT|) (reported beside RESIP)."""
        out = np.zeros(3)
        capi.check(self.L.fvm_divergence(self.h, capi._dptr(out)), "divergence")
        return {"max": float(out[0]), "mean": float(out[1]), "rel": float(out[2])}

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


class Ensemble:
    """Independent cases stepped together by fvm_ensemble_step (one stream per case, nthreads
    host threads)."""

    def __init__(self, solvers, nthreads=8):
        self.L = capi.load()
        self.solvers = list(solvers)
        self.nthreads = int(nthreads)
        self._h = (ctypes.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])

    def step(self):
        n = len(self.solvers)
        infos = (capi.StepInfo * n)()
        code = self.L.fvm_ensemble_step(self._h, n, self.nthreads, infos)
        if code not in (0, 4):
            raise capi.FvmError(code, "fvm_ensemble_step")
        return [infos[q].as_dict() for q in range(n)]

    def sync(self):
        for s in self.solvers:
            s.sync()


def ensemble2d(cases, nthreads=8, dt_scale=1.0, **kw):
    """An Ensemble of Solver2D built from fvm_amd.grid.ensemble_cases entries, each at its
    case's time step times dt_scale."""
    solvers = []
    for c in cases:
        s = Solver2D(c["ni"], c["nj"], c["re"], c["dt"] * dt_scale, c["xg"][:, :, 0],
                     c["yg"][:, :, 0], **kw)
        ic = c["ic"]
        s.set_state(*(ic[k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly")))
        solvers.append(s)
    return Ensemble(solvers, nthreads)


class Solver2D(Solver3D):
    """2D fractional 4-step solver (one-cell span slab of the 3D kernels, see fvm_create2d).
    Fields are (NI+1, NJ+1) arrays of the single plane, as in the 2D reference's RESTART.DAT."""

    def __init__(self, ni, nj, re, dtime, xg, yg, nswpp=20, resimax=1e-4, poisson_mode=2,
                 check_every=1, af_mode=0, af_tol=1e-14):
        self.L = capi.load()
        self.ni, self.nj, self.nk = ni, nj, 2
        self.shape = (ni + 1, nj + 1)
        self.prm = capi.Params(ni, nj, 2, float(re), float(dtime), 1.0, 0, int(nswpp),
                               float(resimax), int(poisson_mode), int(check_every),
                               int(af_mode), float(af_tol))
        flat = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape((ni, nj), order="F")
                                     .ravel(order="F")) for a in (xg, yg)]
        h = ctypes.c_void_p()
        capi.check(self.L.fvm_create2d(ctypes.byref(self.prm), *[capi._dptr(a) for a in flat],
                                       ctypes.byref(h)), "fvm_create2d")
        self.h = h

    def set_state(self, u, v, p, nlx, nly):
        z = np.zeros(self.shape)
        for n, a in zip(("u", "v", "w", "p", "nlx", "nly", "nlz"), (u, v, z, p, nlx, nly, z)):
            self.set(n, a)


class Post2D:
    """Post-processing phase of the 2D CALCCOEF (calccoef.f90:156-195): the Strouhal number
    from successive sign changes of dC_L, the C_L amplitude, running means of C_D, C_L, St,
    C_pb and the rms of C_D and C_L about their running means."""

    def __init__(self, cl_last=0.0):
        # cl_last: C_L of the step before the post phase (CALCCOEF keeps CL_NM1 every step)
        self.toggle = 1
        self.time1 = self.time2 = 0.0
        self.cl_prev = cl_last
        self.cl1 = 0.0
        self.st = 0.0
        self.cl_amp = 0.0
        self.n = 0
        self.cd_ta = self.cl_ta = self.st_ta = self.cpb_ta = self.amp_ta = 0.0
        self.rms_cd = self.rms_cl = 0.0

    def update(self, rtime, cd, cl, cpb):
        self.n += 1
        if float(self.toggle) * (cl - self.cl_prev) < 0.0:
            self.toggle = -self.toggle
            self.st = 1.0 / (rtime - self.time2)
            self.time2 = self.time1
            self.time1 = rtime
            self.cl_amp = abs(cl - self.cl1) / 2.0
            self.cl1 = cl
        self.cl_prev = cl
        self.cd_ta += cd
        self.cl_ta += cl
        self.st_ta += self.st
        self.cpb_ta += cpb
        self.amp_ta += self.cl_amp
        mcd, mcl = self.cd_ta / self.n, self.cl_ta / self.n
        self.rms_cd += (cd - mcd) ** 2
        self.rms_cl += (cl - mcl) ** 2
        return dict(cd_mean=mcd, cl_mean=mcl, st=self.st, st_mean=self.st_ta / self.n,
                    cl_amp=self.cl_amp, cpb_mean=self.cpb_ta / self.n,
                    cd_rms=math.sqrt(self.rms_cd / self.n), cl_rms=math.sqrt(self.rms_cl / self.n))
