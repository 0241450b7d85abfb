"""Mixed-precision forms against the FP64 form over a shedding window.

python -m fvm_amd.precision_report [--workload les3d_re3900] [--spinup 10000] [--steps 10000]
python -m fvm_amd.precision_report --state tests/golden/r2d_65x129_re100_shedding.npz --steps 5600

3D: the LES is started from its seeded field (20 steps at dt/10) and spun up in the FP64 form
for --spinup steps at dt; from that developed state each precision form runs --steps steps --
fp64 (FP64 separable Poisson inverse, FP64 AF sweeps), mixed-poisson (the separable inverse in
FP32 inside the FP64 defect correction) and mixed (that, plus the I and J AF stages solved for
their change in FP32 against the FP64 residual), all to the same scaled-L1 Poisson residual --
and one JSON line reports the end-of-window relative differences of u, v, w, p against fp64
(max-norm over the field maximum), the largest C_L difference over the window and the window's
mean C_D, rms C_L, C_L amplitude and Strouhal number (C_L zero crossings) of every form.

2D (--state): a 2D case from a developed, shedding state (the trajectory fixture's inputs) on the
batched 2D solver in the same three forms (mixed: the FP32 Poisson corrections and the AF sweeps
of both stages in the FP32 change form), all to the same RESIMAX, with the same differences and
statistics.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


FORMS3D = {"fp64": (2, 0), "mixed-poisson": (3, 0), "mixed": (3, 2)}
# 2D batch flags per form: (FP32 Poisson corrections, FP32 change form of the AF sweeps)
FORMS2D = {"fp64": (False, False), "mixed-poisson": (True, False), "mixed": (True, True)}


def solver3d(w, dt, poisson_mode, af_mode, resimax):
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    (xg, yg, zg), ic = make_case(**w)
    s = Solver3D(w["ni"], w["nj"], w["nk"], w["re"], dt, xg, yg, zg, ref_a=math.pi,
                 use_les=True, nswpp=50, resimax=resimax, poisson_mode=poisson_mode,
                 check_every=8, af_mode=af_mode)
    return s, ic


STATE3D = ("u", "v", "w", "p", "nlx", "nly", "nlz")


def spinup3d(w, steps, resimax, startup, log):
    s, ic = solver3d(w, w["dt"] * 0.1, 2, 0, resimax)
    s.set_state(*(ic[k] for k in STATE3D))
    for _ in range(startup):
        s.step()
    s.set_dtime(w["dt"])
    for q in range(steps):
        info = s.step()
        if (q + 1) % 1000 == 0:
            log(f"spin-up: step {q + 1}/{steps} cd {info['cd']:.5f} cl {info['cl']:.3e} "
                f"cfl {info['cfl']:.3f} resip {info['resip']:.2e}")
    state = {k: s.get(k) for k in STATE3D}
    s.close()
    return state


def window3d(w, state, form, steps, resimax, log):
    pm, am = FORMS3D[form]
    s, _ = solver3d(w, w["dt"], pm, am, resimax)
    s.set_state(*(state[k] for k in STATE3D))
    cds, cls, res = [], [], []
    t0 = time.perf_counter()
    for q in range(steps):
        info = s.step()
        cds.append(info["cd"])
        cls.append(info["cl"])
        res.append(info["resip"])
        if (q + 1) % 1000 == 0:
            log(f"{form}: step {q + 1}/{steps} cd {info['cd']:.5f} cl {info['cl']:.3e} "
                f"cfl {info['cfl']:.3f} resip {info['resip']:.2e} passes {info['niter']}")
    s.sync()
    secs = time.perf_counter() - t0
    fields = {c: s.get(c) for c in ("u", "v", "w", "p")}
    s.close()
    cd, cl = np.array(cds), np.array(cls)
    t = w["dt"] * np.arange(1, steps + 1)
    stats = {"cd_mean": float(cd.mean()), "cl_rms": float(np.sqrt(np.mean((cl - cl.mean()) ** 2))),
             "cl_amp": float(0.5 * (cl.max() - cl.min())), "st": crossings_st(t, cl),
             "resip_max": float(max(res)), "secs": secs}
    log(f"{form}: {stats}")
    return fields, stats, cl


def main3d(a, log):
    from fvm_amd.grid import WORKLOADS
    w = dict(WORKLOADS[a.workload])
    if a.dt > 0:
        w["dt"] = a.dt
    state = spinup3d(w, a.spinup, a.resimax, a.startup, log)
    res = {f: window3d(w, state, f, a.steps, a.resimax, log) for f in FORMS3D}
    out = {"workload": a.workload, "dt": w["dt"], "spinup": a.spinup, "steps": a.steps,
           "resimax": a.resimax}
    out.update(compare(res))
    print(json.dumps(out), flush=True)


def crossings_st(t, cl):
    """Strouhal number from the upward zero crossings of C_L (linear interpolation)."""
    up = np.where((cl[:-1] < 0.0) & (cl[1:] >= 0.0))[0]
    if len(up) < 2:
        return float("nan")
    tc = t[up] - cl[up] * (t[up + 1] - t[up]) / (cl[up + 1] - cl[up])
    return float(1.0 / np.mean(np.diff(tc)))


def run2d(g, form, steps, resimax, log):
    from fvm_amd.batch2d import Batch2D
    ni, nj = int(g["ni"]), int(g["nj"])
    mixed, af32 = FORMS2D[form]
    b = Batch2D([dict(ni=ni, nj=nj, re=float(g["re"]), dt=float(g["dt"]), xg=g["xg"],
                      yg=g["yg"])], mixed=mixed, resimax=resimax, npass=6, af32=af32)
    b.set_state(0, *(g["in_" + k] for k in ("u", "v", "p", "nlx", "nly")))
    t0 = time.perf_counter()
    hist = []
    done = 0
    while done < steps:  # the history ring holds 4096 steps
        m = min(2000, steps - done)
        b.step(m)
        hist.append(b.history(m))
        done += m
    secs = time.perf_counter() - t0
    h = np.concatenate(hist)[:, 0, :]
    t = float(g["rtime0"]) + float(g["dt"]) * np.arange(1, steps + 1)
    cd, cl = h[:, 0], h[:, 1]
    stats = {"cd_mean": float(cd.mean()), "cl_rms": float(np.sqrt(np.mean((cl - cl.mean()) ** 2))),
             "cl_amp": float(0.5 * (cl.max() - cl.min())), "st": crossings_st(t, cl),
             "resip_max": float(h[:, 6].max()), "secs": secs}
    log(f"2D {form}: {stats}")
    fields = {c: b.get(0, c) for c in ("u", "v", "p")}
    b.close()
    return fields, stats, cl


def compare(res, base="fp64"):
    """Per form other than base: its statistics and their differences to base's."""
    f64, s64, cl64 = res[base]
    out = {base: s64}
    for f, (fx, sx, clx) in res.items():
        if f == base:
            continue
        rd = lambda k: abs(sx[k] - s64[k]) / abs(s64[k])
        out[f] = dict(sx, field_rel_diff_end={c: float(np.abs(fx[c] - f64[c]).max() /
                                                       np.abs(f64[c]).max()) for c in f64},
                      cl_max_abs_diff=float(np.abs(clx - cl64).max()),
                      cd_mean_rel_diff=rd("cd_mean"), cl_rms_rel_diff=rd("cl_rms"),
                      cl_amp_rel_diff=rd("cl_amp"), st_rel_diff=rd("st"))
    return out


def main2d(a, log):
    g = np.load(a.state)
    res = {f: run2d(g, f, a.steps, a.resimax, log) for f in FORMS2D}
    out = {"state": os.path.basename(a.state), "steps": a.steps, "resimax": a.resimax,
           "dt": float(g["dt"]), "re": float(g["re"]), "grid": [int(g["ni"]), int(g["nj"])]}
    out.update(compare(res))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--state", default="",
                    help="2D: an .npz of a developed state (the trajectory fixture)")
    ap.add_argument("--workload", default="les3d_re3900")
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--spinup", type=int, default=10000, help="3D: FP64 steps before the window")
    ap.add_argument("--resimax", type=float, default=1e-8)
    ap.add_argument("--startup", type=int, default=20)
    ap.add_argument("--dt", type=float, default=0.0, help="override the workload's dt")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.state:
        main2d(a, log)
        return
    main3d(a, log)


if __name__ == "__main__":
    main()
