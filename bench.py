"""Throughput of the 3D dynamic-Smagorinsky LES time step on one MI355X.

Workload: the Re=3900 cylinder LES on a 256 (circumferential) x 192 (radial) x 64 (span) cell
O-grid (NI=193, NJ=257, NK=65 nodes), radial stretching from 0.001 D to 25 D, span pi D,
dt = 0.002, seeded N(0, 0.05) start.  A step is one full fractional 4-step time step (explicit
terms, dynamic SGS model, AF momentum, pressure Poisson solved to the scaled-L1 residual
RESIP <= 1e-8, projection, forces).  The metric is interior cells x steps / wall seconds / 1e6
for the mixed-precision form (FP32 Poisson sweeps in FP64 defect correction) and the plain FP64
form at equal residual; inputs are resident in HBM before the timed region.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12  # bytes/s, MI355X


def build_solver(w, mode, resimax, nswpp, check_every, af_mode=0):
    from fvm_amd.grid import make_case
    from fvm_amd.solver import Solver3D
    (xg, yg, zg), ic = make_case(**w)
    s = Solver3D(w["ni"], w["nj"], w["nk"], w["re"], w["dt"], xg, yg, zg, ref_a=math.pi,
                 use_les=True, nswpp=nswpp, resimax=resimax, poisson_mode=mode,
                 check_every=check_every, af_mode=af_mode)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    return s, ic


def time_mode(w, mode, steps, warmup, resimax, nswpp, check_every, log, af_mode=0):
    s, _ = build_solver(w, mode, resimax, nswpp, check_every, af_mode)
    for q in range(warmup):
        info = s.step()
        log(f"warmup step {q + 1}/{warmup} mode {mode}: poisson passes {info['niter']} "
            f"resip {info['resip']:.2e} af passes {info['af_iters']} cfl {info['cfl']:.3f} "
            f"cd {info['cd']:.4f}")
    s.sync()
    infos = []
    t0 = time.perf_counter()
    for q in range(steps):
        try:
            infos.append(s.step())
        except Exception as e:
            log(f"step {q + 1} mode {mode} failed: {e}")
            raise
    s.sync()
    t1 = time.perf_counter()
    for q, info in enumerate(infos):
        log(f"timed step {q + 1}/{steps} mode {mode}: poisson passes {info['niter']} "
            f"resip {info['resip']:.2e} af passes {info['af_iters']} cfl {info['cfl']:.3f}")
    return s, (t1 - t0), infos


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="les3d_re3900")
    ap.add_argument("--resimax", type=float, default=1e-8)
    ap.add_argument("--nswpp", type=int, default=50)
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--modes", default="mixed,fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="per-kernel HIP-event times")
    ap.add_argument("--af-mode", type=int, default=0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the solver holds the "
              "whole domain on one MI355X; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    from fvm_amd.grid import WORKLOADS
    w = WORKLOADS[a.workload]
    cells = (w["ni"] - 1) * (w["nj"] - 1) * (w["nk"] - 1)
    results = {}
    for mname in a.modes.split(","):
        mode = {"mixed": 3, "fp64": 2, "zebra64": 0, "zebra32": 1}[mname]
        s, secs, infos = time_mode(w, mode, a.steps, a.warmup, a.resimax, a.nswpp,
                                   a.check_every, log, a.af_mode)
        results[mname] = dict(secs=secs, infos=infos, solver=s)
        log(f"{mname}: {secs / a.steps * 1e3:.2f} ms/step "
            f"{cells * a.steps / secs / 1e6:.1f} Mcell-steps/s")
    if a.profile:
        for mname, rr in results.items():
            s = rr["solver"]
            s.profile(True)
            s.profile_reset()
            info = s.step()
            prof = s.profile_get()
            s.profile(False)
            tot = sum(v[0] for v in prof.values())
            log(f"profile {mname}: one step, {info['niter']} outer, kernel total {tot:.2f} ms")
            for k, (ms, calls) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
                log(f"  {k:16s} {ms:9.3f} ms  {calls:6d} calls  {ms / max(calls, 1) * 1e3:9.1f} us/call")
    head = "mixed" if "mixed" in results else list(results)[0]
    r = results[head]
    value = cells * a.steps / r["secs"] / 1e6
    out = {
        "metric": "Mcell-steps/s, 3D LES Re3900, mixed vs FP64 at equal div residual; % HBM peak",
        "value": value, "unit": "Mcell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["secs"] / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": a.workload, "cells": [w["nj"] - 1, w["ni"] - 1, w["nk"] - 1],
                   "re": w["re"], "dt": w["dt"], "resimax": a.resimax,
                   "poisson": "mixed" if head == "mixed" else "fp64"},
    }
    for mname, rr in results.items():
        out[f"{mname}_mcell_steps_per_s"] = cells * a.steps / rr["secs"] / 1e6
        out[f"{mname}_sweeps_per_step"] = float(np.mean([i["niter"] for i in rr["infos"]]))
        out[f"{mname}_resip_max"] = float(max(i["resip"] for i in rr["infos"]))
    if "mixed" in results and "fp64" in results:
        out["mixed_over_fp64"] = results["fp64"]["secs"] / results["mixed"]["secs"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
