"""Batched native 2D solver over the C ABI (fvm_batch2d_*).

Many independent 2D cylinder cases of uneven grid size and Reynolds number (the time step of
cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/, Main.f90:48-118) advanced together: every
phase of the step is one launch over all cases and the whole step is one captured hipGraph.
Fields are (NI+1, NJ+1) arrays in Fortran order, ghosts included, as in the 2D reference's
RESTART.DAT.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import capi

_FIELDS = {"u": 0, "v": 1, "p": 3, "nlx": 4, "nly": 5, "difx": 7, "dify": 8, "s_u": 10,
           "s_v": 11, "djac": 13, "xix": 14, "etx": 15, "xiy": 17, "ety": 18, "x": 23, "y": 24,
           "theta": 26, "un": 31, "vn": 32}

HIST_KEYS = ("cd", "cl", "cd_p", "cd_v", "cl_p", "cl_v", "resip", "cfl")


class Case2D(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("re", ctypes.c_double),
                ("dtime", ctypes.c_double), ("xg", ctypes.POINTER(ctypes.c_double)),
                ("yg", ctypes.POINTER(ctypes.c_double))]


def _bind(L):
    if getattr(L, "_batch2d_bound", False):
        return L
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    pi = ctypes.POINTER(ctypes.c_int)
    L.fvm_batch2d_create.argtypes = [ctypes.POINTER(Case2D), ci, ci, cd, ci, ctypes.POINTER(vp)]
    L.fvm_batch2d_destroy.argtypes = [vp]
    L.fvm_batch2d_set_field.argtypes = [vp, ci, ci, pd]
    L.fvm_batch2d_get_field.argtypes = [vp, ci, ci, pd]
    L.fvm_batch2d_set_dtime.argtypes = [vp, ci, cd]
    L.fvm_batch2d_set_wall_velocity.argtypes = [vp, ci, pd, pd]
    L.fvm_batch2d_step.argtypes = [vp, ci, ctypes.POINTER(capi.StepInfo)]
    L.fvm_batch2d_history.argtypes = [vp, ci, pd]
    L.fvm_batch2d_set_tuning.argtypes = [vp, pi]
    L.fvm_batch2d_get_tuning.argtypes = [vp, pi]
    L.fvm_batch2d_autotune.argtypes = [vp, ci, pi, pd]
    L.fvm_batch2d_profile.argtypes = [vp, ci, ctypes.c_char_p, pd, ctypes.POINTER(ctypes.c_long),
                                      ci]
    L.fvm_batch2d_adjoint_init.argtypes = [vp]
    L.fvm_batch2d_adjoint_step.argtypes = [vp, pd, cd, ci, pd]
    L.fvm_batch2d_adjoint_wall.argtypes = [vp, ci, pd, pd]
    L.fvm_batch2d_adjoint_get.argtypes = [vp, ci, ci, pd]
    L.fvm_batch2d_save.argtypes = [vp, ci]
    L.fvm_batch2d_load.argtypes = [vp, ci]
    for f in ("fvm_batch2d_create", "fvm_batch2d_destroy", "fvm_batch2d_set_field",
              "fvm_batch2d_get_field", "fvm_batch2d_set_dtime", "fvm_batch2d_set_wall_velocity",
              "fvm_batch2d_step", "fvm_batch2d_history", "fvm_batch2d_set_tuning",
              "fvm_batch2d_get_tuning", "fvm_batch2d_autotune", "fvm_batch2d_adjoint_init",
              "fvm_batch2d_adjoint_step", "fvm_batch2d_adjoint_wall", "fvm_batch2d_adjoint_get",
              "fvm_batch2d_save", "fvm_batch2d_load", "fvm_batch2d_profile"):
        getattr(L, f).restype = ci
    L._batch2d_bound = True
    return L


class Batch2D:
    """cases: dicts with ni, nj, re, dt and the (NI, NJ) node arrays xg, yg.  mixed: the Poisson
    corrections in FP32; af32: the AF line sweeps solve for each stage's change in FP32 against
    the stage input's FP64 residual (no adjoint for such a batch)."""

    def __init__(self, cases, mixed=True, resimax=1e-8, npass=3, af32=False):
        self.L = _bind(capi.load())
        self.cases = [dict(c) for c in cases]
        n = len(cases)
        self._nodes = []
        arr = (Case2D * n)()
        for q, c in enumerate(cases):
            xg = np.ascontiguousarray(np.asarray(c["xg"], dtype=np.float64).reshape(
                (c["ni"], c["nj"]), order="F").ravel(order="F"))
            yg = np.ascontiguousarray(np.asarray(c["yg"], dtype=np.float64).reshape(
                (c["ni"], c["nj"]), order="F").ravel(order="F"))
            self._nodes += [xg, yg]
            arr[q] = Case2D(int(c["ni"]), int(c["nj"]), float(c["re"]), float(c["dt"]),
                            capi._dptr(xg), capi._dptr(yg))
        h = ctypes.c_void_p()
        flags = (1 if mixed else 0) | (2 if af32 else 0)
        capi.check(self.L.fvm_batch2d_create(arr, n, flags, float(resimax),
                                             int(npass), ctypes.byref(h)), "fvm_batch2d_create")
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_batch2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, c):
        return (self.cases[c]["ni"] + 1, self.cases[c]["nj"] + 1)

    def set(self, c, name, a):
        a = np.asfortranarray(np.asarray(a, dtype=np.float64).reshape(self.shape(c), order="F"))
        flat = np.ascontiguousarray(a.ravel(order="F"))
        capi.check(self.L.fvm_batch2d_set_field(self.h, c, _FIELDS[name], capi._dptr(flat)),
                   "batch2d set " + name)

    def get(self, c, name):
        flat = np.empty(int(np.prod(self.shape(c))))
        capi.check(self.L.fvm_batch2d_get_field(self.h, c, _FIELDS[name], capi._dptr(flat)),
                   "batch2d get " + name)
        return flat.reshape(self.shape(c), order="F")

    def set_state(self, c, u, v, p, nlx, nly):
        for k, a in zip(("u", "v", "p", "nlx", "nly"), (u, v, p, nlx, nly)):
            self.set(c, k, a)

    def set_dtime(self, c, dt):
        capi.check(self.L.fvm_batch2d_set_dtime(self.h, c, float(dt)), "batch2d set_dtime")
        self.cases[c]["dt"] = float(dt)

    def set_wall_velocity(self, c, uf, vf):
        nj = self.cases[c]["nj"]
        u = np.ascontiguousarray(np.asarray(uf, dtype=np.float64).reshape(nj + 1))
        v = np.ascontiguousarray(np.asarray(vf, dtype=np.float64).reshape(nj + 1))
        capi.check(self.L.fvm_batch2d_set_wall_velocity(self.h, c, capi._dptr(u),
                                                        capi._dptr(v)), "batch2d wall")

    def step(self, nsteps=1):
        """nsteps steps of every case; the last step's scalars per case (status included)."""
        infos = (capi.StepInfo * self.n)()
        code = self.L.fvm_batch2d_step(self.h, int(nsteps), infos)
        if code not in (0, 4):
            raise capi.FvmError(code, "fvm_batch2d_step")
        out = []
        for q in range(self.n):
            d = infos[q].as_dict()
            d["status"] = code
            out.append(d)
        return out

    def history(self, nsteps):
        """(nsteps, ncase, 8) array of C_D, C_L, C_D,p, C_D,v, C_L,p, C_L,v, RESIP, CFL."""
        out = np.empty(nsteps * self.n * 8)
        capi.check(self.L.fvm_batch2d_history(self.h, int(nsteps), capi._dptr(out)),
                   "batch2d history")
        return out.reshape(nsteps, self.n, 8)

    def profile(self, nsteps):
        """nsteps eager steps with HIP events around every launch: {kernel: (total ms, calls)}
        (these steps advance the run like ordinary steps)."""
        cap = 64
        names = ctypes.create_string_buffer(48 * cap)
        ms = np.zeros(cap)
        calls = (ctypes.c_long * cap)()
        q = self.L.fvm_batch2d_profile(self.h, int(nsteps), names, capi._dptr(ms), calls, cap)
        if q < 0:
            raise capi.FvmError(-q, "fvm_batch2d_profile")
        return {names.raw[48 * k:48 * k + 48].split(b"\0")[0].decode(): (float(ms[k]), int(calls[k]))
                for k in range(q)}

    def tuning(self):
        ch = (ctypes.c_int * 2)()
        capi.check(self.L.fvm_batch2d_get_tuning(self.h, ch), "batch2d get_tuning")
        return {"tile_j": ch[0], "fft_cols": ch[1]}

    def set_tuning(self, tile_j, fft_cols):
        ch = (ctypes.c_int * 2)(int(tile_j), int(fft_cols))
        capi.check(self.L.fvm_batch2d_set_tuning(self.h, ch), "batch2d set_tuning")

    def autotune(self, reps=3, cache=None):
        """Time every (stencil tile height, FFT columns) choice on this batch and keep the
        fastest; with cache (a JSON path), a stored choice for the same batch signature is
        applied without timing, and a new choice is stored."""
        sig = self.signature()
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                stored = json.load(fh)
            if sig in stored:
                self.set_tuning(stored[sig]["tile_j"], stored[sig]["fft_cols"])
                return dict(stored[sig], cached=True)
        ch = (ctypes.c_int * 2)()
        ms = (ctypes.c_double * 20)()
        capi.check(self.L.fvm_batch2d_autotune(self.h, int(reps), ch, ms), "batch2d autotune")
        res = {"tile_j": ch[0], "fft_cols": ch[1],
               "ms_per_step": {f"{tj}x{nc}": ms[a * 5 + b] for a, tj in enumerate((1, 2, 4, 8))
                               for b, nc in enumerate((1, 2, 4, 8, 16)) if ms[a * 5 + b] >= 0}}
        if cache:
            stored = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    stored = json.load(fh)
            stored[sig] = res
            with open(cache, "w") as fh:
                json.dump(stored, fh, indent=1, sort_keys=True)
        return res

    # ---- checkpoints and the discrete adjoint ----
    def save(self, slot):
        """Device snapshot of every case's u, v, p, NLX, NLY into slot."""
        capi.check(self.L.fvm_batch2d_save(self.h, int(slot)), "batch2d save")

    def load(self, slot):
        capi.check(self.L.fvm_batch2d_load(self.h, int(slot)), "batch2d load")

    def adjoint_init(self):
        capi.check(self.L.fvm_batch2d_adjoint_init(self.h), "batch2d adjoint_init")

    def adjoint_step(self, weights, tol=1e-13, maxpass=50):
        """One step of the adjoint backwards: the batch must hold the state before the step;
        weights[c] is the weight of case c's C_D after the step.  Returns the worst adjoint
        Poisson residual."""
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(self.n))
        ares = ctypes.c_double()
        capi.check(self.L.fvm_batch2d_adjoint_step(self.h, capi._dptr(w), float(tol),
                                                   int(maxpass), ctypes.byref(ares)),
                   "batch2d adjoint_step")
        return ares.value

    def adjoint_wall(self, c):
        """d(objective)/d(U_F, V_F) of case c per wall cell J (two arrays of NJ+1)."""
        nj = self.cases[c]["nj"]
        gu, gv = np.empty(nj + 1), np.empty(nj + 1)
        capi.check(self.L.fvm_batch2d_adjoint_wall(self.h, c, capi._dptr(gu), capi._dptr(gv)),
                   "batch2d adjoint_wall")
        return gu, gv

    def adjoint_state(self, c, name):
        k = {"u": 0, "v": 1, "p": 2, "nlx": 3, "nly": 4}[name]
        flat = np.empty(int(np.prod(self.shape(c))))
        capi.check(self.L.fvm_batch2d_adjoint_get(self.h, c, k, capi._dptr(flat)),
                   "batch2d adjoint_get")
        return flat.reshape(self.shape(c), order="F")

    def signature(self):
        """The batch's size classes (node counts) and count: the key of a stored tuning."""
        sizes = sorted({(c["ni"], c["nj"]) for c in self.cases})
        return f"n{self.n}:" + ",".join(f"{a}x{b}" for a, b in sizes)


def batch_from_cases(cases, mixed=True, resimax=1e-8, npass=3, dt_scale=1.0, af32=False):
    """A Batch2D of fvm_amd.grid.ensemble_cases entries (grid, seeded start), each at its
    case's time step times dt_scale."""
    specs = [dict(ni=c["ni"], nj=c["nj"], re=c["re"], dt=c["dt"] * dt_scale, xg=c["xg"][:, :, 0],
                  yg=c["yg"][:, :, 0]) for c in cases]
    b = Batch2D(specs, mixed=mixed, resimax=resimax, npass=npass, af32=af32)
    for q, c in enumerate(cases):
        b.set_state(q, *(c["ic"][k][:, :, 0] for k in ("u", "v", "p", "nlx", "nly")))
    return b
