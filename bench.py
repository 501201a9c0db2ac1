#!/usr/bin/env python
"""Throughput of the MI355X-native hot path.

Default workload (--workload cavity): the 3-D lid-driven cavity at Re=400 on a 126^3 hex lattice with
interior vertices jittered by +-20% of the spacing and ~10% of the hexes cut
into prism pairs (about 2.2M polyhedral cells), laminar steady SIMPLE with
least-squares gradients, corrected diffusion, linearUpwind convection;
momentum by BiCGSTAB+Jacobi, pressure correction by CG with the aggregation
AMG, both at the reference cavity case's linear tolerances (relative 1e-8,
absolute 1e-14; cases/cavity/numerics/solution.bs), FP64.  One step is one
SIMPLE outer iteration with every field, matrix and preconditioner resident
in HBM; every timed iteration's linear solves must converge.  The same run
also reports the workload at relative tolerance 1e-6, the mixed-precision
path with its field error against the FP64 run, a short serving sample
(16 requests of 1e4-1e5 cells: mixed served, FP64 served, FP64 one at a
time) and the FP64 oracle on the host cores (median of 3 repeats).

--workload cylinder: transient laminar flow past a circular cylinder (D = 1,
channel 30 D x 20 D, Re = 100) on a structured O-grid + H-grid hex mesh of
about 400k cells extruded one cell; PISO with 2 correctors, BDF2, dt at a
Courant number of 0.5; momentum by FP32 BiCGSTAB inside FP64 refinement,
pressure by FP64 CG + aggregation AMG.  One step is one PISO time step.

--workload channel: steady RANS plane channel with the Wilcox k-omega model
on a 512 x 96 x 100 hex lattice (4.9M cells) graded towards both walls to a
first-cell y+ of about 1, bulk inflow U = 1, Re_bulk = 40,000, inlet k and
omega from 5 % intensity and a 0.1 h length scale; SIMPLE with segregated
omega and k BiCGSTAB in FP32 inside FP64 refinement and the pressure
correction by FP32-AMG CG inside FP64 refinement.  One step is one SIMPLE
outer iteration including both transport solves.

--workload serve: the serving stream of steady laminar lid-driven cavity and
channel requests (256 from a seed, jittered hex meshes log-uniform between
1e4 and 1e6 cells on a ladder of size buckets, Re uniform in [50, 1000],
Poisson arrivals at a fixed rate) through the serving runtime: HBM-budget
admission, concurrent requests on their own streams, resident per-mesh
contexts.  The stream runs with mixed precision (FP32 AMG and inner Krylov
in FP64 refinement) (`--serve-precision both` runs it again in plain FP64;
about 12 minutes per pass); value is the mixed path's converged solves/s,
with p50/p99 time-to-converged.

Prints one JSON line (rank 0, one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

PEAK_HBM_GBS = 8000.0


def measured_traffic(cells):
    """HBM bytes per fine-level sweep launch from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when it was taken on a mesh of
    the same size; None otherwise."""
    path = ROOT / "profiles" / "kJacobiSweep_fine_cavity126_pmc.md"
    if not path.exists():
        return None
    vals = {}
    for line in path.read_text().splitlines():
        if line.startswith("- ") and ":" in line:
            k, v = line[2:].split(":", 1)
            vals[k.strip()] = v.strip().split()[0]
    try:
        if int(vals["cells"]) != cells:
            return None
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def device_memory_used_gb():
    """HBM in use on the current device (hipMemGetInfo), in GB."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return (total.value - free.value) / 1e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["cavity", "cylinder", "channel", "serve"], default="cavity")
    ap.add_argument("--requests", type=int, default=256, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: mean arrivals per second")
    ap.add_argument("--slots", type=int, default=4, help="serve: concurrent requests (worker threads)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serve: device memory budget")
    ap.add_argument("--cells-min", type=float, default=1e4)
    ap.add_argument("--cells-max", type=float, default=1e6)
    ap.add_argument("--serve-precision", choices=["both", "mixed", "fp64"], default="mixed",
                    help="serve: the full stream takes about 12 minutes per precision on one MI355X")
    ap.add_argument("--max-outer", type=int, default=3000, help="serve: outer-iteration cap per request")
    ap.add_argument("--serve-trace", default=None,
                    help="serve: write one CSV row per request (size, iterations, server-clock times) to this path")
    ap.add_argument("--n", type=int, default=126, help="hex lattice cells per side")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--probe-steps", type=int, default=2,
                    help="steps after the timed region in which the dominant kernel is timed with HIP events")
    ap.add_argument("--cpu-iters", type=int, default=None,
                    help="timed oracle iterations for the CPU baseline (3; 1 for the channel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="cpu baseline: timed repeats (median reported)")
    ap.add_argument("--rtol", type=float, default=1e-8,
                    help="cavity: linear relative tolerance of both solves (the reference cavity case: 1e-8)")
    ap.add_argument("--no-legs", action="store_true", help="cavity: skip the other-tolerance and mixed legs")
    ap.add_argument("--serve-sample", type=int, default=16,
                    help="cavity: requests of the short serving sample (0 to skip)")
    ap.add_argument("--rebuild-interval", type=int, default=8)
    ap.add_argument("--amg-strength", type=float, default=0.65,
                    help="cavity: strength threshold of the pressure AMG aggregation")
    ap.add_argument("--amg-correction", type=float, default=2.0,
                    help="cavity: over-correction factor of the pressure AMG coarse-grid correction")
    ap.add_argument("--amg-omega", type=float, default=0.9, help="cavity: damped-Jacobi weight of the AMG smoother")
    ap.add_argument("--mom-pc", choices=["jacobi", "bjacobi", "asm"], default="jacobi",
                    help="cavity: momentum BiCGSTAB preconditioner (bjacobi: ILU(0) in 64-row blocks, the "
                         "reference cavity case's pcType; asm: ILU(0) of the whole matrix)")
    ap.add_argument("--p-pc", choices=["amg", "icc"], default="amg",
                    help="cavity: pressure CG preconditioner (icc: IC(0) of the whole matrix, the reference "
                         "SIMPLE tests' pressure preconditioner)")
    ap.add_argument("--amg-pre", type=int, default=2, help="cavity: AMG pre-smoothing sweeps per level")
    ap.add_argument("--amg-post", type=int, default=2, help="cavity: AMG post-smoothing sweeps per level")
    ap.add_argument("--amg-coarse-size", type=int, default=48, help="cavity: AMG coarsest-level row target")
    ap.add_argument("--warm-start", action="store_true",
                    help="all workloads: start the linear solves from the current unknowns "
                         "(the equation.<name>.warmStart case setting; off, its default, in the cavity and "
                         "cylinder_re100 cases)")
    ap.add_argument("--stream-launches", action="store_true",
                    help="launch the AMG cycles kernel by kernel instead of replaying captured graphs "
                         "(same results; for tracing tools)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 plain 1 GiB device-to-device copies (the known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--read-ceiling", action="store_true",
                    help="only time torch's read-only sum over 256 MiB and 2 GiB of HBM (each after a 1 GiB "
                         "read pass that evicts the Infinity Cache) and exit: the practical read bandwidth "
                         "that the SpMV and smoother figures compare with")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27  # 2^27 doubles = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "copies": 5}), flush=True)
    return 0


def read_ceiling():
    import torch
    flush = torch.zeros(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    out = torch.empty((), dtype=torch.float64, device="cuda")
    res = {}
    for mib in (256, 2048):
        a = torch.ones((mib << 20) // 8, dtype=torch.float64, device="cuda")
        times = []
        for _ in range(12):
            flush.sum()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            torch.sum(a, dim=0, out=out)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) * 1e-3)
        t = sorted(times[2:])[len(times[2:]) // 2]
        res[f"{mib}MiB"] = {"seconds": t, "GB/s": a.numel() * 8 / t / 1e9}
        del a
    print(json.dumps({"read_only_sum_after_eviction": res}), flush=True)
    return 0


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.read_ceiling:
        return read_ceiling()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    if args.stream_launches:
        from babelsim_amd import set_graph_replay
        set_graph_replay(False)
    if args.workload == "cylinder":
        return bench_cylinder(args)
    if args.workload == "channel":
        return bench_channel(args)
    if args.workload == "serve":
        return bench_serve(args)
    import cases
    from babelsim_amd import solver_config
    from babelsim_amd.device import Context, Simple

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    n = args.n
    t0 = time.time()
    mesh = cases.cavity3d_mesh(n, jitter=0.2, split=0.1, seed=20261203)
    nv, nf_in, _, _, nc_in = mesh.sizes()
    log(f"mesh generated: {nc_in} cells, {nf_in} faces ({time.time() - t0:.1f}s)")
    prec = 1 if args.precision == "mixed" else 0

    def config(precision, rtol, mom_pc=None):
        # cases/cavity/numerics/solution.bs:6-16: momentum bcgs and pressure cg
        # at relativeTolerance 1e-8, absoluteTolerance 1e-14.
        common = dict(atol=1e-14, precision=precision, rebuild_interval=args.rebuild_interval,
                      warm_start=int(args.warm_start))
        mpc = {"bjacobi": 3, "asm": 5}.get(mom_pc or args.mom_pc, 1) if precision == 0 else 1
        ppc = 4 if args.p_pc == "icc" and precision == 0 else 2
        return cases.c2_config(
            mom=solver_config(solver=1, pc=mpc, rtol=rtol, **common),
            pre=solver_config(solver=0, pc=ppc, rtol=rtol, amg_strength=args.amg_strength,
                              amg_correction=args.amg_correction, amg_omega=args.amg_omega,
                              amg_pre_sweeps=args.amg_pre, amg_post_sweeps=args.amg_post,
                              amg_coarse_size=args.amg_coarse_size, **common))

    cfg = config(prec, args.rtol)
    ub, pb = cases.cavity_bcs(two_d=False)
    ctx = Context(mesh, renumber=True)
    log(f"context ready: N={ctx.nc} F={ctx.nf} nnz_off={ctx.nnz_off} orthogonal={ctx.orthogonal} "
        f"({time.time() - t0:.1f}s)")

    def run_leg(c, name):
        """warmup + steps outer iterations from rest; every timed iteration's
        linear solves must have converged."""
        sim = Simple(ctx, c, ub, pb)
        sim.start()
        for k in range(args.warmup):
            r = sim.iterate(1)[-1]
            log(f"{name} warmup {k}: rU={r.rU:.3e} mom_its={list(r.mom_iters)} p_its={r.p_iters}")
        before = sim.timing()
        t1 = time.time()
        timed = sim.iterate(args.steps)
        t2 = time.time()
        after = sim.timing()
        bad = [k for k, r in enumerate(timed) if r.numerical_failure or r.mom_status != 0 or r.p_status != 0]
        if bad:
            r = timed[bad[0]]
            raise RuntimeError(f"{name}: timed iteration {bad[0]} has unconverged linear solves "
                               f"(mom_status {r.mom_status}, p_status {r.p_status}, failure {r.numerical_failure})")
        sec = t2 - t1
        last = timed[-1]
        log(f"{name}: {args.steps} its in {sec:.3f}s -> {args.steps / sec:.3f} it/s; last rU={last.rU:.3e} "
            f"mom_its={list(last.mom_iters)} p_its={last.p_iters} mass={last.mass:.2e}")
        timing = {k: (after[k] - before[k]) / args.steps for k in after if k != "iterations"}
        return sim, sec, timed, timing

    sim, sec, timed, timing = run_leg(cfg, f"rtol {args.rtol:g}")
    fields_main = sim.fields()
    its_per_s = args.steps / sec
    last = timed[-1]
    # The timed region replays each AMG V-cycle from a captured graph, where
    # events cannot bracket a single kernel; the sweep is timed with HIP
    # events around every launch in the iterations that follow, which launch
    # the same kernels from the stream.
    sim.probe_start(capacity=512 * args.probe_steps)
    sim.iterate(args.probe_steps)
    launches, sweep_s, sweep_bytes = sim.probe_stop()
    used_gb = device_memory_used_gb()
    # Dominant kernel: the fine-level smoother sweep of the pressure AMG.
    sweep_avg = sweep_s / max(launches, 1)
    achieved = sweep_bytes / sweep_avg / 1e9 if launches else 0.0
    log(f"fine-level sweep: {launches} launches, {sweep_avg * 1e6:.1f} us each, {achieved:.0f} GB/s")
    # Auxiliary: the pressure matrix SpMV alone (after the timed region),
    # from HBM (each launch after a pass that evicts the 256 MB Infinity
    # Cache) and back to back (where the cache may hold part of the matrix).
    spmv_s, N, nnz_off = sim.time_spmv(1, reps=20, cold=True)
    spmv_b2b_s, _, _ = sim.time_spmv(1, reps=50)
    nnz = N + nnz_off
    spmv_bytes = 12 * nnz_off + 24 * N
    log(f"spmv: {spmv_s * 1e6:.1f} us per launch from HBM, {spmv_bytes / spmv_s / 1e9:.0f} GB/s; "
        f"back to back {spmv_b2b_s * 1e6:.1f} us, {spmv_bytes / spmv_b2b_s / 1e9:.0f} GB/s")
    perf_total, perf_systems = sim.performance()
    del sim
    legs = {}
    if not args.no_legs:
        # The same workload at the looser linear tolerance of the earlier
        # measurements, and in mixed precision (FP32 AMG and FP32 BiCGSTAB
        # inside FP64 refinement) with each field's error against the FP64
        # run after the same iterations from rest.
        other = 1e-6 if args.rtol != 1e-6 else 1e-8
        s2, sec2, t2_, _ = run_leg(config(prec, other), f"rtol {other:g}")
        legs[f"rtol_{other:g}"] = {"value": args.steps / sec2, "unit": "outer_iterations/s",
                                   "p_its_last": t2_[-1].p_iters, "mom_its_last": list(t2_[-1].mom_iters)}
        del s2
        if prec == 0:
            s3, sec3, t3_, _ = run_leg(config(1, args.rtol), "mixed")
            Um, pm, phim = s3.fields()
            Uf, pf, phif = fields_main
            rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))  # noqa: E731
            legs["mixed"] = {"value": args.steps / sec3, "unit": "outer_iterations/s",
                             "speedup_over_fp64": (args.steps / sec3) / its_per_s,
                             "p_its_last": t3_[-1].p_iters, "mom_its_last": list(t3_[-1].mom_iters),
                             "field_rel_l2_vs_fp64": {"U": rel(Um, Uf), "p": rel(pm, pf), "phi": rel(phim, phif)},
                             "compared": f"after {args.warmup + args.steps} outer iterations from rest"}
            log(f"mixed: {legs['mixed']['value']:.2f} it/s, field rel L2 vs FP64 {legs['mixed']['field_rel_l2_vs_fp64']}")
            del s3
            # The reference cavity case's momentum preconditioner (pcType bjacobi).
            other_pc = "bjacobi" if args.mom_pc == "jacobi" else "jacobi"
            s4, sec4, t4_, _ = run_leg(config(0, args.rtol, other_pc), f"momentum {other_pc}")
            legs[f"momentum_{other_pc}"] = {"value": args.steps / sec4, "unit": "outer_iterations/s",
                                            "p_its_last": t4_[-1].p_iters, "mom_its_last": list(t4_[-1].mom_iters)}
            del s4
    traffic = measured_traffic(ctx.nc)
    serve = None
    if args.serve_sample > 0:
        serve = serve_sample(args, log)
    cpu = None
    cpu_iters = 1 if args.cpu_iters is None else args.cpu_iters
    if not args.no_cpu and cpu_iters > 0:
        import oracle_lib as orc
        arrays = mesh.arrays()
        t3 = time.time()
        om = orc.oracle_mesh(arrays)
        so = orc.OracleSimple(om, cfg, ub, pb)
        so.start()
        so.iterate(1)  # the first outer iteration is outside the timing window (runtime.cpp:117-130)
        t4 = time.time()
        rates = []
        for _ in range(args.cpu_repeats):
            ta = time.time()
            so.iterate(cpu_iters)
            rates.append(cpu_iters / (time.time() - ta))
        threads = int(orc.oracle().orc_threads())
        rates_sorted = sorted(rates)
        cpu = {"value": float(np.median(rates)), "unit": "outer_iterations/s", "cores": threads, "kind": "port",
               "min": rates_sorted[0], "max": rates_sorted[-1], "repeats": len(rates),
               "host": host_cpu(),
               "sample": f"{len(rates)} repeats of {cpu_iters} FP64 oracle SIMPLE outer iteration(s) on the same "
                         f"{nc_in}-cell mesh at the same linear tolerances, after one untimed iteration (setup "
                         f"{t4 - t3:.1f}s excluded); median reported; OpenMP over {threads} threads, every face "
                         "and row loop parallel"}
        log(f"cpu baseline: median {cpu['value']:.4f} it/s (min {cpu['min']:.4f}, max {cpu['max']:.4f})")
    line = {
        "metric": "SIMPLE outer iterations/s at 2M polyhedral cells (SpMV HBM GB/s in roofline)",
        "value": its_per_s,
        "unit": "outer_iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"3-D lid-driven cavity Re=400, {n}^3 hex lattice, +-20% jitter, ~10% prism splits",
                   "cells": ctx.nc, "faces": ctx.nf, "nnz": nnz, "precision": args.precision,
                   "linear_rtol": args.rtol, "linear_atol": 1e-14,
                   "pressure": (f"CG + aggregation AMG (strength {args.amg_strength}, correction "
                                f"{args.amg_correction}, Jacobi weight {args.amg_omega}, sweeps "
                                f"{args.amg_pre}/{args.amg_post}, coarsest {args.amg_coarse_size} rows)"
                                if args.p_pc == "amg" else "CG + IC(0) of the whole matrix"),
                   "momentum": "BiCGSTAB + " + {"bjacobi": "block Jacobi (ILU(0) per 64 rows)",
                                                "asm": "ILU(0) of the whole matrix"}.get(args.mom_pc, "Jacobi"),
                   "warm_start": args.warm_start,
                   "timed_linear_solves": "all converged",
                   "last_iteration": {"p_its": last.p_iters, "mom_its": list(last.mom_iters)},
                   "phase_seconds_per_iteration": timing, "device_memory_used_GB": used_gb,
                   "performance": {k: perf_total[k] for k in ("linear_solves", "krylov_iterations",
                                                               "preconditioner_setups", "preconditioner_reuses",
                                                               "true_residual_checks", "matrix_value_updates")},
                   "legs": legs, "serve_sample": serve},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": "kJacobiSweep (pressure AMG fine-level smoother sweep, sliced rows)",
                     "launches": launches, "seconds_per_launch": sweep_avg,
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "timing": f"HIP events around every sweep launch over {args.probe_steps} iterations run "
                               "right after the timed region with stream launches",
                     "spmv": {"seconds_per_launch": spmv_s, "algorithmic_bytes_per_launch": spmv_bytes,
                              "GB/s": spmv_bytes / spmv_s / 1e9, "frac": spmv_bytes / spmv_s / 1e9 / PEAK_HBM_GBS,
                              "cache": "each launch after a 512 MiB pass that evicts the Infinity Cache",
                              "back_to_back_GB/s": spmv_bytes / spmv_b2b_s / 1e9}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def host_cpu() -> dict:
    """Model, sockets and cores of the host, and the CPUs this process may use."""
    info = {"usable_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            key = key.strip()
            if key in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[key] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def serve_sample(args, log):
    """A short serving stream in the default run: mixed precision served,
    FP64 served, and the plain FP64 GPU path (one request at a time)."""
    import serve_cases as sc
    specs = sc.stream(count=args.serve_sample, rate=8.0, cells_min=1e4, cells_max=1e5)
    meshes = {}
    for sp in specs:
        key = (sp.kind, sp.side)
        if key not in meshes:
            meshes[key] = sc.mesh_for(sp)
    out = {"requests": len(specs), "cells": [int(min(sp.cells for sp in specs)), int(max(sp.cells for sp in specs))],
           "arrival_rate": 8.0, "outer_tolerance": 1e-5}
    plain = argparse.Namespace(**vars(args))
    plain.slots = 1
    for name, prec, a in (("mixed", 1, args), ("fp64", 0, args), ("fp64_plain_one_at_a_time", 0, plain)):
        r = run_stream(a, specs, meshes, prec, log)
        out[name] = {k: r[k] for k in ("solves_per_s", "converged", "requests", "makespan_s", "p50_s", "p99_s")}
        log(f"serve sample {name}: {r['solves_per_s']:.3f} solves/s, p50 {r['p50_s']:.2f}s, p99 {r['p99_s']:.2f}s")
    out["mixed_over_fp64_served"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["mixed_over_fp64_plain"] = out["mixed"]["solves_per_s"] / out["fp64_plain_one_at_a_time"]["solves_per_s"]
    return out


def bench_cylinder(args):
    import cylinder_mesh as cm
    from babelsim_amd import BDF2, piso_config, solver_config
    from babelsim_amd.device import Context, Piso

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    arrays = cm.cylinder_arrays()
    ctx = Context(arrays, renumber=True)
    log(f"cylinder mesh: N={ctx.nc} F={ctx.nf} nnz_off={ctx.nnz_off} ({time.time() - t0:.1f}s)")
    # cases/cylinder_re100 numerics with two correctors: linearUpwind,
    # least-squares gradients, corrected momentum diffusion, 2 non-orthogonal
    # corrections.  The pressure uses limitedCorrected: on this O+H grid
    # (non-orthogonality up to 33 degrees next to the O-grid square's
    # corners) the uncapped explicit correction of the total-pressure
    # equation grows from step to step.  The pressure AMG aggregates with the
    # strength threshold 0.85: next to the wall the O-grid cells are
    # stretched (first layer 0.005 D), and only their wall-normal couplings
    # count as strong.
    cfg = piso_config(rho=1.0, mu=0.01, m_grad=1, m_diff=1, m_coef_grad=1, p_grad=1, p_diff=2, p_coef_grad=1,
                      time_method=BDF2, n_correctors=2, non_orth_corr=2, tol_mass=1e-7,
                      mom=solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-12, precision=1,
                                        warm_start=int(args.warm_start)),
                      pre=solver_config(solver=0, pc=2, rtol=1e-9, atol=1e-13, amg_strength=0.85,
                                        warm_start=int(args.warm_start)))
    bcs = cm.cylinder_bcs()
    U0 = cm.potential_flow(ctx.geometry()["C"])
    sim = Piso(ctx, cfg, *bcs)
    # dt at Courant 0.5: the Courant number is linear in dt for a given flux.
    sim.set_fields(U0, None)
    sim.start()
    probe_dt = 1e-5
    co = sim.step(probe_dt).co_max / probe_dt
    dt = 0.5 / co
    sim.set_fields(U0, np.zeros(ctx.nc))
    sim.start()
    log(f"dt = {dt:.3e} (Courant 0.5 on the potential-flow start)")
    for k in range(args.warmup):
        r = sim.step(dt)
        log(f"warmup {k}: status={r.status} CoMax={r.co_max:.3f} mom_its={list(r.mom_iters)} p_its={r.p_iters} "
            f"coupling={r.r_coupling:.2e} mass={r.mass:.1e}")
        if r.status != 0:
            raise RuntimeError("PISO warmup step was not accepted")
    before = sim.timing()
    t1 = time.time()
    reps = [sim.step(dt) for _ in range(args.steps)]
    t2 = time.time()
    after = sim.timing()
    sim.probe_start(capacity=4096 * args.probe_steps)
    for _ in range(args.probe_steps):
        sim.step(dt)
    launches, sweep_s, sweep_bytes = sim.probe_stop()
    for k, r in enumerate(reps):
        if r.status != 0:
            raise RuntimeError(f"timed PISO step {k} was not accepted: status={r.status} linear_ok={r.linear_ok} "
                               f"mass={r.mass:.2e} coupling={r.r_coupling:.2e} mom_its={list(r.mom_iters)} "
                               f"p_its={r.p_iters} lin_u={r.lin_u:.2e} lin_p={r.lin_p:.2e}")
    sec = t2 - t1
    last = reps[-1]
    log(f"timed {args.steps} steps in {sec:.3f}s -> {args.steps / sec:.2f} steps/s; CoMax={last.co_max:.3f} "
        f"mom_its={list(last.mom_iters)} p_its={last.p_iters}")
    timing = {k: (after[k] - before[k]) / args.steps for k in after if k != "steps"}
    sweep_avg = sweep_s / max(launches, 1)
    achieved = sweep_bytes / sweep_avg / 1e9 if launches else 0.0
    cpu = None
    cpu_iters = 3 if args.cpu_iters is None else args.cpu_iters
    if not args.no_cpu and cpu_iters > 0:
        import oracle_lib as orc
        t3 = time.time()
        om = orc.oracle_mesh(arrays)
        so = orc.OraclePiso(om, cfg, *bcs)
        so.set_fields(cm.potential_flow(om.geometry()["C"]), np.zeros(om.nc))
        so.start()
        so.step(dt)  # the first step outside the timing window
        t4 = time.time()
        for _ in range(cpu_iters):
            so.step(dt)
        t5 = time.time()
        threads = int(orc.oracle().orc_threads())
        cpu = {"value": cpu_iters / (t5 - t4), "unit": "time_steps/s", "cores": threads, "kind": "port",
               "sample": f"{cpu_iters} timed FP64 oracle PISO steps on the same {ctx.nc}-cell mesh after one "
                         f"untimed step (setup {t4 - t3:.1f}s excluded); OpenMP over {threads} threads"}
        log(f"cpu baseline: {cpu['value']:.4f} steps/s")
    line = {
        "metric": "PISO time steps/s, cylinder Re=100 at 400k hex cells",
        "value": args.steps / sec,
        "unit": "time_steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "cylinder D=1 in a 30Dx20D channel, Re=100, O+H hex grid extruded one cell",
                   "cells": ctx.nc, "faces": ctx.nf, "dt": dt, "correctors": 2, "non_orthogonal_corrections": 2,
                   "pressure_diffusion": "limitedCorrected",
                   "time": "BDF2", "momentum": "FP32 BiCGSTAB + Jacobi in FP64 refinement",
                   "pressure": "FP64 CG + aggregation AMG (strength 0.85)", "warm_start": args.warm_start,
                   "last_step": {"CoMax": last.co_max, "p_its": last.p_iters, "mom_its": list(last.mom_iters),
                                 "coupling": last.r_coupling, "mass": last.mass},
                   "phase_seconds_per_step": timing, "device_memory_used_GB": device_memory_used_gb()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "kJacobiSweep (pressure AMG fine-level smoother sweep, sliced rows)",
                     "launches": launches, "seconds_per_launch": sweep_avg,
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "timing": f"HIP events around every sweep launch over {args.probe_steps} steps run "
                               "right after the timed region with stream launches"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_channel(args):
    import channel_mesh as ch
    from babelsim_amd.device import Context, Simple

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    mesh = ch.graded_channel_mesh()
    ctx = Context(mesh, renumber=True)
    log(f"graded channel: N={ctx.nc} F={ctx.nf} nnz_off={ctx.nnz_off} ({time.time() - t0:.1f}s)")
    cfg, kw = ch.channel_config(precision=1, rebuild_interval=args.rebuild_interval, warm_start=int(args.warm_start))
    ub, pb, kb, wb = ch.channel_bcs()
    sim = Simple(ctx, cfg, ub, pb)
    sim.enable_komega(kw, kb, wb)
    U0, k0, w0 = ch.initial_fields(ctx.nc)
    sim.set_fields(U0, np.zeros(ctx.nc))
    sim.set_turbulence(k0, w0)
    sim.start()
    for k in range(args.warmup):
        r = sim.iterate(1)[0]
        log(f"warmup {k}: rU={r.rU:.3e} mom_its={list(r.mom_iters)} p_its={r.p_iters} k_its={r.k_iters} "
            f"w_its={r.w_iters} dTurb={r.d_turb:.2e}")
    before = sim.timing()
    t1 = time.time()
    timed = sim.iterate(args.steps)
    t2 = time.time()
    after = sim.timing()
    sim.probe_start(capacity=512 * args.probe_steps)
    sim.iterate(args.probe_steps)
    launches, sweep_s, sweep_bytes = sim.probe_stop()
    for k, r in enumerate(timed):
        if not r.turb_linear or r.mom_status != 0 or r.p_status != 0:
            raise RuntimeError(f"timed iteration {k} has an unconverged linear solve: mom_status={r.mom_status} "
                               f"p_status={r.p_status} turb_linear={r.turb_linear}")
    sec = t2 - t1
    last = timed[-1]
    log(f"timed {args.steps} its in {sec:.3f}s -> {args.steps / sec:.3f} it/s; rU={last.rU:.3e} "
        f"mom_its={list(last.mom_iters)} p_its={last.p_iters} k_its={last.k_iters} w_its={last.w_iters}")
    timing = {k: (after[k] - before[k]) / args.steps for k in after if k != "iterations"}
    sweep_avg = sweep_s / max(launches, 1)
    achieved = sweep_bytes / sweep_avg / 1e9 if launches else 0.0
    log(f"fine-level sweep: {launches} launches, {sweep_avg * 1e6:.1f} us each, {achieved:.0f} GB/s")
    cpu = None
    cpu_iters = 1 if args.cpu_iters is None else args.cpu_iters
    if not args.no_cpu and cpu_iters > 0:
        import oracle_lib as orc
        t3 = time.time()
        om = orc.oracle_mesh(mesh.arrays())
        so = orc.OracleSimpleKOmega(om, cfg, kw, ub, pb, kb, wb)
        U0, k0, w0 = ch.initial_fields(om.nc)
        so.set_fields(U0, np.zeros(om.nc))
        so.set_turbulence(k0, w0)
        so.start()
        t4 = time.time()
        so.iterate(cpu_iters)
        t5 = time.time()
        threads = int(orc.oracle().orc_threads())
        cpu = {"value": cpu_iters / (t5 - t4), "unit": "outer_iterations/s", "cores": threads, "kind": "port",
               "sample": f"the first {cpu_iters} FP64 oracle SIMPLE k-omega outer iteration(s) on the same "
                         f"{om.nc}-cell mesh (setup {t4 - t3:.1f}s excluded); OpenMP over {threads} threads, "
                         "the assembly face loops serial"}
        log(f"cpu baseline: {cpu['value']:.4f} it/s")
    line = {
        "metric": "SIMPLE k-omega outer iterations/s, graded plane channel at 4.9M hex cells",
        "value": args.steps / sec,
        "unit": "outer_iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "plane channel 512x96x100 hex, wall grading to y+ ~1, Re_bulk=40000, Wilcox k-omega",
                   "cells": ctx.nc, "faces": ctx.nf, "nnz": ctx.nc + ctx.nnz_off,
                   "momentum": "FP32 BiCGSTAB + Jacobi in FP64 refinement",
                   "turbulence": "omega then k, FP32 BiCGSTAB + Jacobi in FP64 refinement",
                   "pressure": "FP32-AMG CG in FP64 refinement (strength 0.85)", "warm_start": args.warm_start,
                   "last_iteration": {"p_its": last.p_iters, "mom_its": list(last.mom_iters),
                                      "k_its": last.k_iters, "w_its": last.w_iters, "rU": last.rU,
                                      "d_turb": last.d_turb},
                   "phase_seconds_per_iteration": timing, "device_memory_used_GB": device_memory_used_gb()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "kernel": "kJacobiSweep (pressure AMG fine-level smoother sweep, sliced rows)",
                     "launches": launches, "seconds_per_launch": sweep_avg,
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "timing": f"HIP events around every sweep launch over {args.probe_steps} iterations run "
                               "right after the timed region with stream launches"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_stream(args, specs, meshes, precision, log):
    """One pass of the stream through a fresh server; returns its summary."""
    import serve_cases as sc
    from babelsim_amd import SERVE_CONVERGED
    from babelsim_amd.serve import Server

    srv = Server(slots=args.slots, memory_budget=args.budget_gb * 1e9)
    case_ids = {key: srv.add_case(m) for key, m in meshes.items()}
    lead = 0.2
    t0 = srv.clock() + lead
    ids = []
    for sp in specs:
        cfg = sc.serve_config(sp.re, precision=precision, warm_start=int(args.warm_start))
        ub, pb = sc.bcs_for(sp)
        ids.append(srv.submit_steady(case_ids[(sp.kind, sp.side)], cfg, ub, pb, max_iterations=args.max_outer,
                                     arrival=t0 + sp.arrival))
    # Progress lines every ~15 s while the stream drains.
    pending, last_log = set(ids), time.time()
    while pending:
        time.sleep(0.25)
        pending = {rid for rid in pending if not srv.done(rid)}
        if time.time() - last_log > 15.0 or not pending:
            last_log = time.time()
            log(f"{'mixed' if precision else 'fp64'}: {len(ids) - len(pending)}/{len(ids)} done at "
                f"t={srv.clock() - t0:.1f}s")
    results = [srv.wait(rid) for rid in ids]
    stats = srv.stats()
    srv.close()
    if args.serve_trace:
        with open(f"{args.serve_trace}.{'mixed' if precision else 'fp64'}.csv", "w") as f:
            f.write("index,kind,cells,re,status,iterations,slot,reused,released,admitted,started,finished\n")
            for sp, r in zip(specs, results):
                f.write(f"{sp.index},{sp.kind},{sp.cells},{sp.re:.1f},{r.status},{r.iterations},{r.slot},"
                        f"{int(r.reused_context)},{r.released - t0:.4f},{r.admitted - t0:.4f},"
                        f"{r.started - t0:.4f},{r.finished - t0:.4f}\n")
    ok = [r for r in results if r.status == SERVE_CONVERGED]
    end = max(r.finished for r in results)
    makespan = end - t0
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    queue = np.array([r.admitted - r.released for r in results])
    its = np.array([r.iterations for r in results])
    return {
        "solves_per_s": len(ok) / makespan,
        "converged": len(ok), "requests": len(results),
        "statuses": {str(s): int(sum(r.status == s for r in results)) for s in sorted({r.status for r in results})},
        "makespan_s": makespan,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "mean_queue_s": float(queue.mean()),
        "outer_iterations": {"total": int(its.sum()), "mean": float(its.mean()), "max": int(its.max())},
        "cell_iterations_per_s": float(sum(sp.cells * r.iterations for sp, r in zip(specs, results)) / makespan),
        "server": stats,
        "max_measured_solver_bytes": max(r.measured_bytes for r in results),
    }


def bench_serve(args):
    import serve_cases as sc

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    specs = sc.stream(count=args.requests, rate=args.rate, cells_min=args.cells_min, cells_max=args.cells_max)
    meshes = {}
    for sp in specs:
        key = (sp.kind, sp.side)
        if key not in meshes:
            meshes[key] = sc.mesh_for(sp)
    log(f"stream: {len(specs)} requests over {specs[-1].arrival:.1f}s of arrivals, {len(meshes)} distinct meshes, "
        f"{sum(sp.cells for sp in specs)} cells in total ({time.time() - t0:.1f}s)")
    runs = {}
    order = {"both": (1, 0), "mixed": (1,), "fp64": (0,)}[args.serve_precision]
    for prec in order:
        name = "mixed" if prec else "fp64"
        t1 = time.time()
        runs[name] = run_stream(args, specs, meshes, prec, log)
        r = runs[name]
        log(f"{name}: {r['converged']}/{r['requests']} converged in {r['makespan_s']:.1f}s -> "
            f"{r['solves_per_s']:.3f} solves/s, p50 {r['p50_s']:.2f}s p99 {r['p99_s']:.2f}s "
            f"(wall {time.time() - t1:.1f}s)")
    head = runs.get("mixed") or runs["fp64"]
    cpu = None
    cpu_iters = 1 if args.cpu_iters is None else args.cpu_iters
    if not args.no_cpu and cpu_iters > 0:
        # The oracle solves the stream's first request of the smallest
        # bucket to convergence, FP64, on the host cores.
        import oracle_lib as orc
        sp = min(specs, key=lambda q: (q.cells, q.index))
        ub, pb = sc.bcs_for(sp)
        om = orc.oracle_mesh(meshes[(sp.kind, sp.side)].arrays())
        so = orc.OracleSimple(om, sc.serve_config(sp.re, warm_start=int(args.warm_start)), ub, pb)
        so.start()
        t4 = time.time()
        n = 0
        for n in range(1, args.max_outer + 1):
            if so.iterate(1)[-1].converged:
                break
        t5 = time.time()
        threads = int(orc.oracle().orc_threads())
        cpu = {"value": 1.0 / (t5 - t4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"one request of the stream ({sp.kind}, {sp.cells} cells, Re {sp.re:.0f}) solved to "
                         f"convergence ({n} outer iterations) by the FP64 oracle SIMPLE, OpenMP over {threads} "
                         "threads; the smallest request of the stream"}
        log(f"cpu baseline: {cpu['value']:.4f} solves/s on the smallest request")
    line = {
        "metric": "served solves/s with p50/p99 time-to-converged (steady laminar stream, 1e4-1e6 cells)",
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(specs),
        "warmup": 0,
        "ms_per_step": 1e3 / head["solves_per_s"] if head["solves_per_s"] > 0 else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if "mixed" in runs else "f64",
        "data": "synthetic",
        "config": {"workload": f"serving stream: {len(specs)} steady laminar cavity/channel requests, jittered hex "
                               f"meshes log-uniform in [{args.cells_min:.0e}, {args.cells_max:.0e}] cells on "
                               f"{len(meshes)} size buckets, Re ~ U[50, 1000], Poisson arrivals at "
                               f"{args.rate}/s",
                   "slots": args.slots, "budget_GB": args.budget_gb, "outer_tolerance": 1e-5,
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                   "runs": runs,
                   "mixed_over_fp64": (runs["mixed"]["solves_per_s"] / runs["fp64"]["solves_per_s"]
                                       if "mixed" in runs and "fp64" in runs else None)},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
