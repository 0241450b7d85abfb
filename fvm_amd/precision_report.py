"""Mixed-precision form against the FP64 form on the same seeded start.

python -m fvm_amd.precision_report [--workload les3d_re3900_25m] [--steps 500]

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="les3d_re3900_25m")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--resimax", type=float, default=1e-8)
    ap.add_argument("--startup", type=int, default=20)
    ap.add_argument("--dt", type=float, default=0.0, help="override the workload's dt")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
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
