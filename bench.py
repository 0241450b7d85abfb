"""Throughput of the 3D dynamic-Smagorinsky LES time step on one MI355X.

Workload: the Re=3900 cylinder LES on a 256 (circumferential) x 192 (radial) x 64 (span) cell
O-grid (NI=193, NJ=257, NK=65 nodes), radial stretching from 0.001 D to 25 D, span pi D,
dt = 0.002, seeded N(0, 0.05) start.  A step is one full fractional 4-step time step (explicit
terms, dynamic SGS model, AF momentum, pressure Poisson solved to the scaled-L1 residual
RESIP <= 1e-8, projection, forces).  The metric is interior cells x steps / wall seconds / 1e6
for the mixed-precision form (FP32 approximate inverse inside FP64 defect correction) and the
plain FP64 form at equal residual; inputs are resident in HBM before the timed region.

The impulsive cold start at dt = 0.002 diverges in the reference scheme on this grid (the
wall-adjacent radial Courant number of the start field is 2), so the untimed start-up runs
--startup-steps steps at dt/10 before switching to dt = 0.002 (the reference reassigns DTIME at
run time the same way).  Per-kernel durations are HIP events on the solver's stream over as many
steps again right after the timed steps; the dominant kernel's algorithmic bytes per launch over
its mean duration gives the roofline fraction against the 8 TB/s HBM peak.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per interior cell of one launch: each distinct cell-array operand of the
# kernel's dataflow read once, each output written once, FP64 = 8 B (FP32 = 4 B); ghosts and
# O(surface) work excluded.  'units' is how many cell sweeps one launch performs.
KERNEL_BYTES = {
    # one pass of the line-scan AF stage: line factors {beta, 1/D, B, c_(a-1)} shared by the
    # three components, {Q} -> {x} per component: 32 + 3 x 16 bytes
    "k_scan_pass_i": (80, 1), "k_scan_pass_j": (80, 1),
    # exact K-line sweep: Thomas factors {A, D, B} shared, {SU} -> {x} per component
    "k_af_thomas_k": (72, 1),
    # chunked fallback pass over the 3 components: {phi_in, SU, J, 3 stage metric products}
    # -> phi_out = 7 doubles per cell and component
    "k_af_chunk_i": (56, 3), "k_af_chunk_j": (56, 3), "k_af_chunk_k": (56, 3),
    # The metric sets below count 7 of the 9 metrics: xi_z and eta_z vanish identically on the
    # extruded grid and are not read.
    # strain rates fused with the 27-point test filter: {u,v,w, xi 7} -> {S 6, |S|, u^ 3}
    "k_strain_filter": (8 * 20, 1),
    # AF right-hand side, stage 1: {phi 3 (= U_N at the first stage of a step), 6 metric
    # products, 2 Re J/dt, NL 3, NL_NM1 3, DIF 3, S 3, 3 face coefficients} -> SU 3; stage 2:
    # {phi 3, products 6, 2 Re J/dt, faces 3} -> SU 3
    "k_af_rhs_i": (8 * 28, 1), "k_af_rhs_j": (8 * 16, 1),
    # Poisson residual (mixed form): {p, rhs, D_E, D_N, D_T, J^-1 Q12, Q13, Q23} -> r (FP32)
    "k_residual": (8 * 8 + 4, 1),
    # DIF: {u_n, v_n, w_n, D_E, D_N, D_T, J^-1 Q12, Q13, Q23} -> DIF 3
    "dif": (8 * 12, 1),
    # P_SOURCE + projection: {p, J, xi 7, u, v, w} -> {S 3, u, v, w}
    "psource_correct": (8 * 18, 1),
    # Germano identity terms: {u,v,w, u^ 3, xi 7, S 6, Delta^2} -> workgroup sums of LM, MM
    # (|S| is formed from S as it is staged)
    "k_lm": (8 * 20, 1),
    # SGS divergence in the step: {xi 7, nu_t, S 6, J, NL 3} -> {NL 3} (SGS_x,y,z are formed
    # when they are read)
    "k_sgs_div": (8 * 21, 1),
}


def cells_of(w):
    return (w["ni"] - 1) * (w["nj"] - 1) * (w["nk"] - 1)


def build_solver(w, mode, resimax, nswpp, check_every, dt, af_mode=0, thermal=False):
    from fvm_amd.grid import make_case, seeded_temperature
    from fvm_amd.solver import Solver3D
    (xg, yg, zg), ic = make_case(**w)
    s = Solver3D(w["ni"], w["nj"], w["nk"], w["re"], dt, xg, yg, zg, ref_a=math.pi,
                 use_les=True, nswpp=nswpp, resimax=resimax, poisson_mode=mode,
                 check_every=check_every, af_mode=af_mode)
    s.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    if thermal:  # the passive temperature at the reference's Peclet number
        t = seeded_temperature(w["ni"], w["nj"], w["nk"], 0.05)
        s.set_temperature(t, 0.0 * t, pe=2769.0)
    return s


def run_mode(w, mode, a, log):
    dt = w["dt"]
    s = build_solver(w, mode, a.resimax, a.nswpp, a.check_every, dt * a.startup_factor,
                     a.af_mode, a.thermal)
    for q in range(a.startup_steps):
        info = s.step()
        if q == 0 or q + 1 == a.startup_steps:
            log(f"start-up step {q + 1}/{a.startup_steps} (dt {dt * a.startup_factor:g}): "
                f"poisson passes {info['niter']} cfl {info['cfl']:.3f} cd {info['cd']:.3f}")
    s.set_dtime(dt)
    for q in range(a.warmup):
        info = s.step()
        log(f"warmup step {q + 1}/{a.warmup} mode {mode}: poisson passes {info['niter']} "
            f"resip {info['resip']:.2e} af passes {info['af_iters']} cfl {info['cfl']:.3f} "
            f"cd {info['cd']:.4f}")
    s.sync()
    infos = []
    t0 = time.perf_counter()
    for q in range(a.steps):
        infos.append(s.step())
    s.sync()
    t1 = time.perf_counter()
    # the same number of steps again with HIP events around every kernel on the solver's
    # stream (kept out of the timed region: recording events adds host work per launch)
    s.profile(True)
    s.profile_reset()
    for q in range(a.steps):
        s.step()
    s.sync()
    prof = s.profile_get()
    s.profile(False)
    for q, info in enumerate(infos):
        log(f"timed step {q + 1}/{a.steps} mode {mode}: poisson passes {info['niter']} "
            f"resip {info['resip']:.2e} af passes {info['af_iters']} cfl {info['cfl']:.3f} "
            f"cd {info['cd']:.4f}")
    return dict(secs=t1 - t0, infos=infos, prof=prof)


def roofline(prof, ncell):
    rows = []
    for name, (b, units) in KERNEL_BYTES.items():
        if name not in prof or prof[name][1] == 0:
            continue
        ms, calls = prof[name]
        avg_s = ms / calls * 1e-3
        nbytes = b * units * ncell
        rows.append(dict(kernel=name, calls=calls, avg_us=avg_s * 1e6, total_ms=ms,
                         bytes_per_launch=nbytes, achieved_gbs=nbytes / avg_s / 1e9,
                         frac=nbytes / avg_s / 1e9 / HBM_PEAK_GBS))
    rows.sort(key=lambda r: -r["total_ms"])
    return rows


def traffic_for(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a
    copy-kernel calibration (profiles/traffic.json), or None when not measured."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def cpu_baseline(w, log):
    """The oracle (C restatement of the reference's step, one thread) on one C1 time step with
    the reference's own sweep cap NSWPP = 20, after one untimed step."""
    from fvm_amd.grid import make_case
    from oracle.oracle3d import Oracle3D
    (xg, yg, zg), ic = make_case(**w)
    o = Oracle3D(w["ni"], w["nj"], w["nk"], w["re"], w["dt"] * 0.1, math.pi, xg, yg, zg)
    o.set_state(*(ic[k] for k in ("u", "v", "w", "p", "nlx", "nly", "nlz")))
    o.step(1, 1e-8, True)
    t0 = time.perf_counter()
    o.step(20, 1e-8, True)
    t1 = time.perf_counter()
    log(f"cpu oracle: one step with 20 Poisson sweeps in {t1 - t0:.2f} s")
    return dict(value=cells_of(w) / (t1 - t0) / 1e6, unit="Mcell-steps/s", cores=1, kind="port",
                sample="one 192x256x64 time step (SGS, AF, 20 lexicographic Poisson line "
                       "sweeps = the reference's NSWPP, RESIP not driven to 1e-8), oracle C "
                       f"restatement on 1 host core, {t1 - t0:.2f} s")


def copy_calibration(mib, log):
    """A plain device copy of a known size (torch's copy kernel, four times): under rocprofv3
    --pmc FETCH_SIZE (or WRITE_SIZE) each copy dispatch reads (writes) exactly mib MiB, which
    fixes the scale factor applied to the counters of the solver's kernels."""
    import torch
    n = mib * 1024 * 1024 // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB x 4")


def run_ensemble(a, log):
    """The ensemble of seeded 2D cases (uneven sizes, Re in [40, 300]) stepped together by
    fvm_ensemble_step; the same steps with one host thread give the one-case-at-a-time rate."""
    from fvm_amd.grid import ensemble_cases
    from fvm_amd.solver import ensemble2d
    cases = ensemble_cases(a.cases)
    ncell = sum((c["ni"] - 1) * (c["nj"] - 1) for c in cases)
    log(f"ensemble: {len(cases)} cases, {ncell} cells, {a.threads} host threads")
    rates = {}
    for nt in (a.threads, 1):
        # the impulsive start (uniform flow through the wall cells) at dt/10 first, as for
        # the 3D workloads, then each case's own dt
        ens = ensemble2d(cases, nthreads=nt, dt_scale=a.startup_factor, nswpp=a.nswpp,
                         resimax=a.resimax, poisson_mode=3, check_every=a.check_every)
        for q in range(a.startup_steps):
            infos = ens.step()
        log(f"start-up: {a.startup_steps} steps at dt/{1 / a.startup_factor:g}, max cfl "
            f"{max(i['cfl'] for i in infos):.3f}")
        for sv, c in zip(ens.solvers, cases):
            sv.set_dtime(c["dt"])
        for q in range(a.warmup):
            infos = ens.step()
            log(f"warmup step {q + 1}/{a.warmup} threads {nt}: max cfl "
                f"{max(i['cfl'] for i in infos):.3f} max resip {max(i['resip'] for i in infos):.2e}")
        ens.sync()
        t0 = time.perf_counter()
        for q in range(a.steps):
            infos = ens.step()
        ens.sync()
        secs = time.perf_counter() - t0
        assert all(np.isfinite(i["cd"]) for i in infos)
        rates[nt] = (secs, infos)
        log(f"threads {nt}: {secs / a.steps * 1e3:.2f} ms/step "
            f"{ncell * a.steps / secs / 1e6:.1f} Mcell-steps/s")
        del ens
    secs, infos = rates[a.threads]
    out = {
        "metric": "Mcell-steps/s, ensemble of 2D cylinder cases of uneven size and Re",
        "value": ncell * a.steps / secs / 1e6, "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": secs / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "ensemble2d", "cases": len(cases), "cells": ncell,
                   "host_threads": a.threads, "resimax": a.resimax, "poisson": "mixed",
                   "startup": f"{a.startup_steps} untimed steps at dt/{1 / a.startup_factor:g}"},
        "one_thread_mcell_steps_per_s": ncell * a.steps / rates[1][0] / 1e6,
        "resip_max": float(max(i["resip"] for i in infos)),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="les3d_re3900")
    ap.add_argument("--resimax", type=float, default=1e-8)
    ap.add_argument("--nswpp", type=int, default=50)
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--modes", default="mixed,fp64")
    ap.add_argument("--startup-steps", type=int, default=20)
    ap.add_argument("--startup-factor", type=float, default=0.1)
    ap.add_argument("--af-mode", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="print every kernel scope")
    ap.add_argument("--thermal", action="store_true",
                    help="advance the passive temperature too (not part of the headline step)")
    ap.add_argument("--cases", type=int, default=96, help="ensemble2d: number of 2D cases")
    ap.add_argument("--threads", type=int, default=8, help="ensemble2d: host threads")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB four times (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the solver holds "
              "the whole domain on one MI355X (the span-slab sharding is a design note only)",
              file=sys.stderr)
        sys.exit(2)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration, log)
        return
    if a.workload == "ensemble2d":
        run_ensemble(a, log)
        return
    from fvm_amd.grid import WORKLOADS
    w = WORKLOADS[a.workload]
    ncell = cells_of(w)
    results = {}
    for mname in a.modes.split(","):
        mode = {"mixed": 3, "fp64": 2, "zebra64": 0, "zebra32": 1}[mname]
        r = run_mode(w, mode, a, log)
        results[mname] = r
        log(f"{mname}: {r['secs'] / a.steps * 1e3:.2f} ms/step "
            f"{ncell * a.steps / r['secs'] / 1e6:.1f} Mcell-steps/s")
        if a.profile:
            tot = sum(v[0] for k, v in r["prof"].items() if not k.startswith("k_"))
            log(f"profile {mname}: {a.steps} steps, scope total {tot:.2f} ms")
            for k, (ms, calls) in sorted(r["prof"].items(), key=lambda kv: -kv[1][0]):
                log(f"  {k:16s} {ms:9.3f} ms  {calls:6d} calls  "
                    f"{ms / max(calls, 1) * 1e3:9.1f} us/call")
    head = "mixed" if "mixed" in results else list(results)[0]
    r = results[head]
    value = ncell * a.steps / r["secs"] / 1e6
    rows = roofline(r["prof"], ncell)
    roof = None
    if rows:
        top = rows[0]
        roof = {"bound": "hbm", "achieved": top["achieved_gbs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": top["frac"], "traffic": traffic_for(top["kernel"]),
                "kernel": top["kernel"], "bytes_per_launch": top["bytes_per_launch"],
                "avg_us": top["avg_us"], "launches": top["calls"]}
    out = {
        "metric": "Mcell-steps/s, 3D LES Re3900, mixed vs FP64 at equal div residual; % HBM peak",
        "value": value, "unit": "Mcell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["secs"] / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": a.workload, "cells": [w["nj"] - 1, w["ni"] - 1, w["nk"] - 1],
                   "re": w["re"], "dt": w["dt"], "resimax": a.resimax,
                   "poisson": "mixed" if head == "mixed" else "fp64",
                   "startup": f"{a.startup_steps} untimed steps at dt={w['dt'] * a.startup_factor:g}",
                   "passive_temperature": bool(a.thermal)},
        "roofline": roof,
    }
    for mname, rr in results.items():
        out[f"{mname}_mcell_steps_per_s"] = ncell * a.steps / rr["secs"] / 1e6
        out[f"{mname}_poisson_passes_per_step"] = float(np.mean([i["niter"] for i in rr["infos"]]))
        out[f"{mname}_resip_max"] = float(max(i["resip"] for i in rr["infos"]))
    if "mixed" in results and "fp64" in results:
        out["mixed_over_fp64"] = results["fp64"]["secs"] / results["mixed"]["secs"]
    out["kernels"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in row.items()}
                      for row in rows]
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline(w, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
