#!/usr/bin/env python3
"""Benchmark: accepted simulated seconds per wall second of the coupled
free-surface step at 310x48x80 (the reference beach, water/lab_runner.gd:26,
plain method at the reference's PCG tolerance) with the seeded 1e-3 m
alongshore perturbation, over the 20 simulated seconds of shoaling, breaking,
runup and backwash.

A "step" is one accepted solver step (CFL substep trial accepted on the
device).  W warmup steps run untimed from the reset state, then exactly K
accepted steps are timed, bracketed by device synchronisation.  Batches are
submitted ahead with one control readback per batch (water/work_scheduler.gd),
the quota stops the device after step K.

Also reported: the standalone pressure solve at 310x48x80 (seeded wavy
surface, N(0, 0.1 m/s) faces, zero-mean RHS) in milliseconds per solve at
relative residual 1e-6 and 1e-8, plain FP32 MGPCG against the mixed form
(BF16 V-cycle, FP64 dots and true residual), the roofline of the dominant kernel (level-0 Chebyshev sweep
of the multigrid preconditioner, timed live with HIP events on the solver's
stream) and the CPU oracle timed on the host as a baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from waves_amd.settings import benchmark_beach, default_beach  # noqa: E402
from waves_amd.solver import Solver, stopped  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "accepted sim-s per wall-s @310x48x80; pressure ms/step @1e-6 rel residual"


def run_steps(solver: Solver, count: int, ahead: int = 2, batches_per_submit: int = 1):
    """Advance `count` more accepted steps; returns (control, batches submitted)."""
    c = solver.poll(wait=True)
    target = int(c[2]) + count
    solver.request(target)
    submitted = 0
    inflight = 0
    while True:
        while inflight < ahead:
            solver.frame(batches_per_submit)
            submitted += batches_per_submit
            inflight += 1
        c = solver.poll(wait=True)  # waits for the newest readback
        inflight = 0
        if c[2] >= target or stopped(c):
            return c, submitted


def cpu_baseline(sample_steps: int = 1):
    """The CPU oracle (a scalar restatement, one host thread) on the same
    workload: reset, then `sample_steps` accepted steps."""
    from oracle.oracle import Oracle
    s = benchmark_beach()
    o = Oracle(s)
    o.initial()
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        o.step()
    wall = time.perf_counter() - t0
    c = o.control
    return {"value": float(c[1]) / wall, "unit": "accepted sim-s per wall-s", "cores": 1, "kind": "port",
            "sample": f"oracle, {sample_steps} accepted step(s) from reset at 310x48x80 ({float(c[1]):.4f} sim-s in {wall:.1f} s)"}


def pressure_leg(repeats: int = 3):
    """Standalone solves on the pressure case at 310x48x80: per tolerance and
    form, one untimed solve then `repeats` solves from a fresh load; the
    device time (HIP events, hierarchy audit .. projection) is averaged."""
    from waves_amd.scenarios import load_pressure_case
    s = default_beach(pressure_tolerance=1e-6)
    g = Solver(s)
    g.start()
    out = {"workload": "standalone pressure solve, 310x48x80, seeded wavy surface, N(0,0.1 m/s) faces, zero-mean RHS"}
    for tol in (1e-6, 1e-8):
        for mode, name in ((0, "plain_fp32"), (1, "mixed_bf16_fp64")):
            load_pressure_case(g, s)
            g.pressure_solve(mode, tol)
            runs = []
            for _ in range(repeats):
                load_pressure_case(g, s)
                runs.append(g.pressure_solve(mode, tol))
            ms = float(np.mean([r["device_ms"] for r in runs]))
            r = runs[-1]
            out[f"{name}@{tol:g}"] = {"ms": ms, "iterations": r["iterations"], "restarts": r["restarts"],
                                      "converged": r["converged"], "true_rel_residual": r["true_rel_residual"]}
        out[f"speedup@{tol:g}"] = out[f"plain_fp32@{tol:g}"]["ms"] / out[f"mixed_bf16_fp64@{tol:g}"]["ms"]
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-pressure", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)

    s = benchmark_beach()
    solver = Solver(s)
    solver.start()
    c, _ = run_steps(solver, args.warmup)
    solver.synchronize()
    t_start_sim, steps_start, iters_start, rej_start = float(c[1]), int(c[2]), float(c[18]), float(c[3])
    t0 = time.perf_counter()
    c, batches = run_steps(solver, args.steps)
    solver.synchronize()
    wall = time.perf_counter() - t0
    c = solver.poll(wait=True)
    sim = float(c[1]) - t_start_sim
    steps = int(c[2]) - steps_start
    value = sim / wall

    # roofline of the dominant kernel: level-0 Chebyshev sweep, algorithmic bytes
    # per cell = reads b, x, active, 6 coefficients, l1 D^-1, d (11 x 4 B) + writes x', d (2 x 4 B)
    N = s.cells
    cheb_us = solver.time_kernel(0, reps=200)
    copy_us = solver.time_kernel(3, reps=200)
    cheb_bytes = 52 * N
    achieved = cheb_bytes / (cheb_us * 1e-6) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_mg_cheb_level0.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "accepted sim-s per wall-s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / max(steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: reference beach from rest, seeded 1e-3 m alongshore perturbation",
        "config": {"workload": "coupled free-surface step, 310x48x80 beach (1,190,400 cells), PCG rel tol 1e-4, 20 sim-s run",
                   "grid": [310, 48, 80], "sim_seconds_timed": sim, "sim_time_start": t_start_sim,
                   "batches": batches, "pcg_iterations": float(c[18]) - iters_start,
                   "rejections": float(c[3]) - rej_start, "failure": int(c[7])},
        "roofline": {"bound": "hbm", "kernel": "mg level-0 Chebyshev sweep", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "avg_us": cheb_us, "bytes_per_launch": cheb_bytes,
                     "copy_GBs": 8 * N / (copy_us * 1e-6) / 1e9},
    }
    solver.close()
    if not args.no_pressure:
        result["pressure"] = pressure_leg()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
