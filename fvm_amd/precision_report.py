"""Mixed-precision form against the FP64 form on the same seeded start.

python -m fvm_amd.precision_report [--workload les3d_re3900_25m] [--steps 500]
python -m fvm_amd.precision_report --state tests/golden/r2d_65x129_re100_shedding.npz --steps 5600

The second form runs a 2D case from a developed, shedding state (the trajectory fixture's
inputs) on the batched 2D solver, FP64 Poisson correction against the FP32 one inside the FP64
defect correction, both to the same RESIMAX, and reports the same differences plus the C_L
amplitude and the Strouhal number from the C_L zero crossings of the window.

Runs the LES twice from the same seeded start -- the FP64 Poisson inverse (poisson_mode 2) and
the FP32 inverse inside FP64 defect correction (poisson_mode 3), both to the same scaled-L1
residual -- through a start-up of 20 steps at dt/10 and then --steps steps at dt, and prints one
JSON line with the relative differences of u, v, w, p (max-norm over the field maximum) and of
the force-coefficient statistics (mean C_D, rms C_L, Strouhal number from the C_L extrema as in
the 2D reference's post-processing).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


def run(w, mode, steps, resimax, startup, log):
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Post2D, Solver3D
    (xg, yg, zg), ic = make_case(**w)
    s = Solver3D(w["ni"], w["nj"], w["nk"], w["re"], w["dt"] * 0.1, xg, yg, zg, ref_a=math.pi,
                 use_les=True, nswpp=50, resimax=resimax, poisson_mode=mode, check_every=8)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    for _ in range(startup):
        s.step()
    s.set_dtime(w["dt"])
    post = Post2D()
    cds, cls = [], []
    t = 0.0
    t0 = time.perf_counter()
    out = None
    for q in range(steps):
        info = s.step()
        t += w["dt"]
        cds.append(info["cd"])
        cls.append(info["cl"])
        out = post.update(t, info["cd"], info["cl"], info["cpb"])
        if (q + 1) % 25 == 0:
            log(f"mode {mode}: step {q + 1}/{steps} cd {info['cd']:.5f} cl {info['cl']:.3e} "
                f"cfl {info['cfl']:.3f} resip {info['resip']:.2e} passes {info['niter']}")
    s.sync()
    secs = time.perf_counter() - t0
    fields = {c: s.get(c) for c in ("u", "v", "w", "p")}
    cl = np.array(cls)
    stats = {"cd_mean": float(np.mean(cds)), "cl_rms": float(np.sqrt(np.mean((cl - cl.mean()) ** 2))),
             "st": out["st"], "secs": secs}
    s.close()
    return fields, stats


def crossings_st(t, cl):
    """Strouhal number from the upward zero crossings of C_L (linear interpolation)."""
    up = np.where((cl[:-1] < 0.0) & (cl[1:] >= 0.0))[0]
    if len(up) < 2:
        return float("nan")
    tc = t[up] - cl[up] * (t[up + 1] - t[up]) / (cl[up + 1] - cl[up])
    return float(1.0 / np.mean(np.diff(tc)))


def run2d(g, mixed, steps, resimax, log):
    from fvm_amd.batch2d import Batch2D
    ni, nj = int(g["ni"]), int(g["nj"])
    b = Batch2D([dict(ni=ni, nj=nj, re=float(g["re"]), dt=float(g["dt"]), xg=g["xg"],
                      yg=g["yg"])], mixed=mixed, resimax=resimax, npass=6)
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
    log(f"2D {'mixed' if mixed else 'fp64'}: {stats}")
    fields = {c: b.get(0, c) for c in ("u", "v", "p")}
    b.close()
    return fields, stats, cl


def main2d(a, log):
    g = np.load(a.state)
    f64, s64, cl64 = run2d(g, False, a.steps, a.resimax, log)
    f32, s32, cl32 = run2d(g, True, a.steps, a.resimax, log)
    rel = {c: float(np.abs(f32[c] - f64[c]).max() / np.abs(f64[c]).max()) for c in f64}
    rd = lambda k: abs(s32[k] - s64[k]) / abs(s64[k])
    out = {"state": os.path.basename(a.state), "steps": a.steps, "resimax": a.resimax,
           "dt": float(g["dt"]), "re": float(g["re"]), "grid": [int(g["ni"]), int(g["nj"])],
           "field_rel_diff_end": rel, "cl_max_abs_diff": float(np.abs(cl32 - cl64).max()),
           "fp64": s64, "mixed": s32, "cd_mean_rel_diff": rd("cd_mean"),
           "cl_rms_rel_diff": rd("cl_rms"), "st_rel_diff": rd("st"), "cl_amp_rel_diff": rd("cl_amp")}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--state", default="",
                    help="2D: an .npz of a developed state (the trajectory fixture)")
    ap.add_argument("--workload", default="les3d_re3900_25m")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--resimax", type=float, default=1e-8)
    ap.add_argument("--startup", type=int, default=20)
    ap.add_argument("--dt", type=float, default=0.0, help="override the workload's dt")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.state:
        main2d(a, log)
        return
    from fvm_amd.grid import WORKLOADS
    w = dict(WORKLOADS[a.workload])
    if a.dt > 0:
        w["dt"] = a.dt
    f64, s64 = run(w, 2, a.steps, a.resimax, a.startup, log)
    f32, s32 = run(w, 3, a.steps, a.resimax, a.startup, log)
    rel = {c: float(np.abs(f32[c] - f64[c]).max() / np.abs(f64[c]).max()) for c in f64}
    out = {"workload": a.workload, "dt": w["dt"], "steps": a.steps, "resimax": a.resimax,
           "field_rel_diff": rel,
           "fp64": s64, "mixed": s32,
           "cd_mean_rel_diff": abs(s32["cd_mean"] - s64["cd_mean"]) / abs(s64["cd_mean"]),
           "cl_rms_rel_diff": abs(s32["cl_rms"] - s64["cl_rms"]) / max(abs(s64["cl_rms"]), 1e-300),
           "st_diff": s32["st"] - s64["st"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
