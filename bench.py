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
    # the same pass in the FP32 change form: FP32 factors 16 bytes shared, {r} -> {delta}
    # 3 x 8 bytes
    "k_scan_pass_i/f32": (40, 1), "k_scan_pass_j/f32": (40, 1),
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
    # NL and DIF in one pass (the step): {u_n, v_n, w_n, JU 3, D_E, D_N, D_T, J^-1 Q12, Q13,
    # Q23} -> {NL 3, DIF 3}
    "nl_dif": (8 * 18, 1),
    # P_SOURCE + projection: {p, J, xi 7, u, v, w} -> {S 3, u, v, w}
    "psource_correct": (8 * 18, 1),
    # Germano identity terms: {u,v,w, u^ 3, xi 7, S 6, Delta^2} -> workgroup sums of LM, MM
    # (|S| is formed from S as it is staged)
    "k_lm": (8 * 20, 1),
    # SGS divergence in the step: {xi 7, nu_t, S 6, J, NL 3} -> {NL 3} (SGS_x,y,z are formed
    # when they are read)
    "k_sgs_div": (8 * 21, 1),
}


# precision forms of the 3D step: (poisson_mode, af_mode; None: --af-mode).  mixed: FP32
# separable Poisson corrections and FP32 changes of the I and J AF stages, each inside FP64
# defect correction; mixed-poisson: the FP32 Poisson corrections only
MODES = {"mixed": (3, 2), "mixed-poisson": (3, None), "fp64": (2, None), "zebra64": (0, None),
         "zebra32": (1, None)}


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


def run_mode(w, mode, a, log, af_mode=None):
    dt = w["dt"]
    s = build_solver(w, mode, a.resimax, a.nswpp, a.check_every, dt * a.startup_factor,
                     a.af_mode if af_mode is None else af_mode, a.thermal)
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
    div = s.divergence()  # the face-flux divergence of the last step's velocity, beside RESIP
    for q, info in enumerate(infos):
        log(f"timed step {q + 1}/{a.steps} mode {mode}: poisson passes {info['niter']} "
            f"resip {info['resip']:.2e} af passes {info['af_iters']} cfl {info['cfl']:.3f} "
            f"cd {info['cd']:.4f}")
    log(f"mode {mode}: face-flux divergence of the last step's velocity: max {div['max']:.3e} "
        f"mean {div['mean']:.3e} relative {div['rel']:.3e}")
    return dict(secs=t1 - t0, infos=infos, prof=prof, div=div)


def roofline(prof, ncell, af32=False):
    rows = []
    for key, (b, units) in KERNEL_BYTES.items():
        name, _, form = key.partition("/")
        if (form == "f32") != (af32 and name.startswith("k_scan_pass")):
            continue
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


REF_MPI = os.path.join(HERE, "oracle", "_ref", "rles_c1_mpi16", "rles")
MPIRUN = "/opt/conda/bin/mpirun"


def cpu_baseline_reference(w, log, nranks=16, nsteps=4, timeout_s=420):
    """The reference's own program on the host cores: LES/v2.3_LES-Reynolds stresses built by
    oracle/ref/build_ref_mpi.sh (its main.f90 and MPI (J, K) block decomposition unchanged,
    4 x 4 ranks) for the 192 x 256 x 64 grid, cold start at dt 2e-4, the reference's own Poisson
    settings (NSWPP = 20 line-GS sweeps per step, RESIMAX 1e-4), run with mpirun -np 16.  The
    step time is the reference's own per-step report (CPU_TIME of rank 0, main.f90:155-158),
    median over the steps after the first; the run is stopped after nsteps steps (before the
    reference's end-of-run output).  None when the build or mpirun is absent or the run fails."""
    import re
    import signal
    import subprocess
    import tempfile
    if not (os.path.exists(REF_MPI) and os.path.exists(MPIRUN)):
        return None
    from fvm_amd.grid import OGrid, grid_filename, write_grid3d
    g = OGrid(w["ni"], w["nj"], w["nk"], w["h1"], w["r_out"], math.pi)
    secs = []
    with tempfile.TemporaryDirectory() as d:
        write_grid3d(os.path.join(d, grid_filename(w["ni"], w["nj"], w["nk"])), *g.nodes())
        t0 = time.perf_counter()
        proc = subprocess.Popen([MPIRUN, "-np", str(nranks), REF_MPI], cwd=d,
                                stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                                stderr=subprocess.STDOUT, text=True, start_new_session=True)
        try:
            for line in proc.stdout:
                m = re.search(r"Time of operation =\s*([0-9.]+)", line)
                if m:
                    secs.append(float(m.group(1)))
                    log(f"cpu reference step {len(secs)}: {secs[-1]:.3f} s")
                if len(secs) >= nsteps or time.perf_counter() - t0 > timeout_s:
                    break
        finally:
            if proc.poll() is None:
                os.killpg(proc.pid, signal.SIGTERM)  # the process group started above
                try:
                    proc.wait(timeout=30)
                except subprocess.TimeoutExpired:
                    os.killpg(proc.pid, signal.SIGKILL)
                    proc.wait()
        wall = time.perf_counter() - t0
    if len(secs) < 2:
        log(f"cpu reference: {len(secs)} steps reported")
        return None
    step = float(np.median(secs[1:]))
    log(f"cpu reference: {len(secs)} steps on {nranks} ranks, wall {wall:.1f} s")
    return dict(value=cells_of(w) / step / 1e6, unit="Mcell-steps/s", cores=nranks,
                kind="reference",
                sample=f"{len(secs)} steps of the reference program (flang build, mpirun -np "
                       f"{nranks}, NDIMJ x NDIMK = 4 x 4) on the 192x256x64 grid from its cold "
                       f"start at dt 2e-4 with its own Poisson settings (20 line-GS sweeps per "
                       f"step, RESIMAX 1e-4: not the 1e-8 of the GPU figure); median of the "
                       f"reference's per-step times after the first, {step:.3f} s")


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


# Algorithmic bytes per interior cell of the batched 2D kernels, counted as KERNEL_BYTES (each
# distinct operand once, FP64 8 B, FP32 4 B; line factors shared by the u and v waves once).
KERNEL_BYTES_2D = {
    # phase A: {u, v, p, J, xi_x, xi_y, eta_x, eta_y, D_E, D_N, J Q12, 2 Re J/dt, stage-1 cross
    # weights, AN, NLX, NLY} -> {U_N, V_N, NL 2, DIF 2, S 2, SU 2}
    "k2_explicit": 8 * 26,
    # I-line sweep: factors {1/D, beta, B, AN} + per component {SU, x_old, x}
    "k2_sweep_i": 8 * 10,
    # J-line sweep: factors {1/D, beta, B, AE, AW} + per component {SU, x_old, x}
    "k2_sweep_j": 8 * 11,
    # stage-2 right-hand side: {u, v, 2 Re J/dt, cross weights, AE} -> SU 2
    "k2_rhs2": 8 * 7,
    # predictor correction: {u, v, S 2, J} -> {u, v}
    "k2_correct": 8 * 7,
    # Poisson source: {u, v, J, 4 metrics} -> RHS
    "k2_prhs": 8 * 8,
    # residual: {p, RHS, D_E, D_N, J Q12} -> r (FP32 in the mixed form)
    "k2_residual": 8 * 5 + 4,
    # projection: {p, J, 4 metrics, u, v} -> {u, v, S 2}
    "k2_project": 8 * 12,
}
# the FP32 change form of the AF sweeps (--ens-form mixed): phase A and the stage-2
# right-hand side read AE / AN instead of 2 Re J/dt and store the residual in FP32; the sweeps
# move the same operands in FP32
KERNEL_BYTES_2D_AF32 = dict(KERNEL_BYTES_2D, k2_explicit=8 * 24 + 4 * 2, k2_sweep_i=4 * 10,
                            k2_sweep_j=4 * 11, k2_rhs2=8 * 5 + 4 * 2)


def cpu_baseline_ensemble(cases, log, budget_s=20.0):
    """The oracle (C restatement of the step on a one-cell span slab, W = 0, Poisson line
    Gauss-Seidel to RESIP 1e-8) on a sample of the ensemble's cases, one host core per worker
    process as the reference runs one case per process: per case one untimed and one timed step;
    the rate is the sum of the per-case rates over the worker count."""
    import concurrent.futures as cf
    nproc = max(1, min(16, os.cpu_count() or 1))
    sample = cases[:nproc]
    t0 = time.perf_counter()
    with cf.ProcessPoolExecutor(nproc) as ex:
        res = list(ex.map(_oracle_case_step, [(c["ni"], c["nj"], c["re"], c["dt"], c["xg"],
                                               c["yg"], c["ic"]) for c in sample]))
    wall = time.perf_counter() - t0
    cells = sum(r[0] for r in res)
    secs = sum(r[1] for r in res)
    log(f"cpu oracle ensemble sample: {len(sample)} cases on {nproc} processes, step time "
        f"{secs:.2f} s summed, wall {wall:.1f} s, sweeps {min(r[2] for r in res)}-"
        f"{max(r[2] for r in res)}, RESIP max {max(r[3] for r in res):.1e}")
    return dict(value=cells / secs * nproc / 1e6, unit="Mcell-steps/s", cores=nproc,
                kind="port",
                sample=f"{len(sample)} of the ensemble's cases, one step each after one untimed "
                       f"step, oracle C restatement (one-cell span slab, Poisson line GS to "
                       f"RESIP 1e-8, cap 20000 sweeps), {nproc} worker processes of one core; "
                       f"per-case step times summed {secs:.1f} s")


def _oracle_case_step(args):
    ni, nj, re, dt, xg, yg, ic = args
    from oracle.oracle3d import Oracle3D
    lz = 1.0e6
    slab = lambda a, k: np.asfortranarray(np.repeat(np.asarray(a)[:, :, :1], k, axis=2))
    zg = np.zeros((ni, nj, 2), order="F")
    zg[:, :, 1] = lz
    o = Oracle3D(ni, nj, 2, re, dt, lz, slab(xg, 2), slab(yg, 2), zg)
    o.set_no_pref(True)
    z = np.zeros((ni + 1, nj + 1, 3), order="F")
    o.set_state(slab(ic["u"], 3), slab(ic["v"], 3), z, slab(ic["p"], 3), slab(ic["nlx"], 3),
                slab(ic["nly"], 3), z)
    o.step(20000, 1e-8, False)
    t0 = time.perf_counter()
    o.step(20000, 1e-8, False)
    secs = time.perf_counter() - t0
    return (ni - 1) * (nj - 1), secs, int(o.scalar("niter")), float(o.scalar("resip"))


def run_ensemble(a, log):
    """The ensemble of seeded 2D cases (uneven sizes, Re in [40, 300]) as one batch
    (fvm_batch2d_*): one launch per phase over all cases, the step captured as a hipGraph and
    replayed, tiles chosen by the batch's autotuner.  Start-up at dt/10, then the time step."""
    from fvm_amd.batch2d import batch_from_cases
    from fvm_amd.grid import ensemble_cases
    cases = ensemble_cases(a.cases, dt=a.ens_dt, dt_rule="uniform")
    ncell = sum((c["ni"] - 1) * (c["nj"] - 1) for c in cases)
    log(f"ensemble: {len(cases)} cases, {ncell} cells, dt {a.ens_dt}")
    b = batch_from_cases(cases, mixed=a.ens_form != "fp64", resimax=a.resimax, npass=a.npass,
                         dt_scale=a.startup_factor, af32=a.ens_form == "mixed")
    infos = b.step(a.startup_steps)
    log(f"start-up: {a.startup_steps} steps at dt/{1 / a.startup_factor:g}, max cfl "
        f"{max(i['cfl'] for i in infos):.3f}")
    for q, c in enumerate(cases):
        b.set_dtime(q, c["dt"])
    if a.ens_tuning:  # a fixed tile choice instead of the autotuner's timed trials
        tj, nc = (int(x) for x in a.ens_tuning.split(","))
        b.set_tuning(tj, nc)
        tune = {"tile_j": tj, "fft_cols": nc, "ms_per_step": {}}
        log(f"tuning given: tile height {tj}, FFT columns {nc}")
    else:
        tune = b.autotune(reps=3)
        log(f"autotuner: tile height {tune['tile_j']}, FFT columns {tune['fft_cols']}; "
            + ", ".join(f"{k} {v:.3f} ms" for k, v in sorted(tune["ms_per_step"].items())))
    for q in range(a.warmup):
        infos = b.step(1)
        log(f"warmup step {q + 1}/{a.warmup}: max cfl {max(i['cfl'] for i in infos):.3f} "
            f"max resip {max(i['resip'] for i in infos):.2e} max passes "
            f"{max(i['niter'] for i in infos)}")
    b.step(0)  # synchronizes the stream
    t0 = time.perf_counter()
    infos = b.step(a.steps)  # a.steps graph replays, one synchronization at the end
    secs = time.perf_counter() - t0
    hist = b.history(a.steps)
    assert np.isfinite(hist).all(), "non-finite C_D / C_L in the timed steps"
    log(f"batch: {secs / a.steps * 1e3:.3f} ms/step {ncell * a.steps / secs / 1e6:.1f} "
        f"Mcell-steps/s; RESIP max {hist[:, :, 6].max():.2e}, CFL max {hist[:, :, 7].max():.3f}")
    prof = b.profile(a.steps)
    tot = sum(v[0] for v in prof.values())
    for k, (ms, calls) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:16s} {ms:9.3f} ms {calls:6d} launches {ms / calls * 1e3:9.1f} us/launch "
            f"({100 * ms / tot:.1f}%)")
    rows = []
    for k, nb in (KERNEL_BYTES_2D_AF32 if a.ens_form == "mixed" else KERNEL_BYTES_2D).items():
        if k in prof:
            ms, calls = prof[k]
            avg = ms / calls * 1e-3
            rows.append(dict(kernel=k, calls=calls, avg_us=avg * 1e6, total_ms=ms,
                             bytes_per_launch=nb * ncell, achieved_gbs=nb * ncell / avg / 1e9,
                             frac=nb * ncell / avg / 1e9 / HBM_PEAK_GBS))
    rows.sort(key=lambda r: -r["total_ms"])
    top = rows[0]
    out = {
        "metric": "Mcell-steps/s, batched ensemble of 2D cylinder cases of uneven size and Re",
        "value": ncell * a.steps / secs / 1e6, "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": secs / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.ens_form == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": "ensemble2d", "form": a.ens_form, "cases": len(cases), "cells": ncell,
                   "dt": a.ens_dt, "resimax": a.resimax, "poisson_passes": a.npass,
                   "poisson": "fp64" if a.ens_form == "fp64" else "mixed",
                   "af": "fp32-change" if a.ens_form == "mixed" else "fp64",
                   "tuning": {"tile_j": tune["tile_j"],
                                                  "fft_cols": tune["fft_cols"]},
                   "startup": f"{a.startup_steps} untimed steps at dt/{1 / a.startup_factor:g}"},
        "resip_max": float(hist[:, :, 6].max()),
        "roofline": {"bound": "hbm", "achieved": top["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": top["frac"], "traffic": None,
                     "kernel": top["kernel"], "bytes_per_launch": top["bytes_per_launch"],
                     "avg_us": top["avg_us"], "launches": top["calls"]},
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                    for r in rows],
        "profile_share": {k: round(v[0] / tot, 4) for k, v in prof.items()},
    }
    out["cpu_baseline"] = None if a.no_cpu else cpu_baseline_ensemble(cases, log)
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
    ap.add_argument("--npass", type=int, default=3,
                    help="ensemble2d: Poisson correction passes per step")
    ap.add_argument("--ens-dt", type=float, default=0.001,
                    help="ensemble2d: the time step of every case (the reference scheme diverges "
                         "on some of the drawn cases at 0.0015 and above)")
    ap.add_argument("--ens-form", default="mixed", choices=("mixed", "mixed-poisson", "fp64"),
                    help="ensemble2d: mixed: FP32 Poisson corrections and the AF sweeps in the "
                         "FP32 change form; mixed-poisson: the Poisson corrections only; fp64: "
                         "neither")
    ap.add_argument("--ens-tuning", default="", metavar="TJ,NCOL",
                    help="ensemble2d: use this tile choice (stencil tile height, FFT columns) "
                         "instead of the autotuner (profiled runs)")
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
        # precision forms: Poisson form (poisson_mode), AF form (af_mode; None: --af-mode)
        mode, afm = MODES[mname]
        r = run_mode(w, mode, a, log, afm)
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
    rows = roofline(r["prof"], ncell, af32=MODES[head][1] == 2)
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
        "dtype": "f64+f32" if head.startswith("mixed") else "f64", "data": "synthetic",
        "config": {"workload": a.workload, "cells": [w["nj"] - 1, w["ni"] - 1, w["nk"] - 1],
                   "re": w["re"], "dt": w["dt"], "resimax": a.resimax,
                   "poisson": "mixed" if MODES[head][0] in (1, 3) else "fp64",
                   "af": "fp32 change form" if MODES[head][1] == 2 else "fp64",
                   "startup": f"{a.startup_steps} untimed steps at dt={w['dt'] * a.startup_factor:g}",
                   "passive_temperature": bool(a.thermal)},
        "roofline": roof,
    }
    for mname, rr in results.items():
        out[f"{mname}_mcell_steps_per_s"] = ncell * a.steps / rr["secs"] / 1e6
        out[f"{mname}_poisson_passes_per_step"] = float(np.mean([i["niter"] for i in rr["infos"]]))
        out[f"{mname}_resip_max"] = float(max(i["resip"] for i in rr["infos"]))
        out[f"{mname}_div_norm"] = rr["div"]
    if "mixed" in results and "fp64" in results:
        out["mixed_over_fp64"] = results["fp64"]["secs"] / results["mixed"]["secs"]
    out["kernels"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in row.items()}
                      for row in rows]
    if a.no_cpu:
        out["cpu_baseline"] = None
    else:
        out["cpu_baseline"] = cpu_baseline_reference(w, log) or cpu_baseline(w, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
