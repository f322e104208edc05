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
(BF16 V-cycle, FP64 dots and true residual) and the mixed form with the
learned warm start; the coupled step with the mixed PCG in the batch program;
the refined 620x96x160 beach; rooflines of the dominant kernel (the coarse
Chebyshev sweep at multigrid level 1) and of the MacCormack trace and the
level-0 sweep, timed live with HIP events on the solver's stream; and the CPU
oracle timed on the host as a baseline.
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


def pressure_leg(repeats: int = 3, n=(310, 48, 80), spacing=(0.025, 0.025, 0.05), learned: bool = True):
    """Standalone solves on the pressure case: per tolerance and form, one
    untimed solve then `repeats` solves from a fresh load; the device time
    (HIP events, hierarchy audit .. projection) is averaged.  Forms: plain
    FP32 MGPCG (the reference's ops and criteria), mixed (BF16 V-cycle, FP64
    dots, FP64 true residual with refinement) and mixed + learned warm start
    (the committed network, waves_amd/data/warmstart_unet.npy)."""
    from waves_amd.scenarios import load_pressure_case
    from waves_amd.learned import WEIGHTS
    s = default_beach(pressure_tolerance=1e-6, n=n, spacing=spacing)
    g = Solver(s)
    g.start()
    forms = [(0, "plain_fp32"), (1, "mixed_bf16_fp64")]
    net = None
    if learned and os.path.exists(WEIGHTS):
        from waves_amd.net import Net
        net = Net(s.n, 1)
        net.set_params(np.load(WEIGHTS))
        g.attach_net(net)
        forms.append((2, "mixed_learned"))
    out = {"workload": f"standalone pressure solve, {n[0]}x{n[1]}x{n[2]}, seeded wavy surface, N(0,0.1 m/s) faces, zero-mean RHS, from p = 0"}
    for tol in (1e-6, 1e-8):
        for mode, name in forms:
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
        out[f"speedup_mixed@{tol:g}"] = out[f"plain_fp32@{tol:g}"]["ms"] / out[f"mixed_bf16_fp64@{tol:g}"]["ms"]
        if net is not None:
            out[f"speedup_learned@{tol:g}"] = out[f"plain_fp32@{tol:g}"]["ms"] / out[f"mixed_learned@{tol:g}"]["ms"]
    g.close()
    if net is not None:
        net.close()
    return out


def inloop_leg(steps: int = 600):
    """The coupled step with the batch program's PCG in plain FP32 and in the
    mixed form (wv_set_precision), same configuration and start: wall ms per
    accepted step and the acceptance record of both."""
    out = {}
    for prec, name in ((0, "plain_fp32"), (1, "mixed_bf16_fp64")):
        g = Solver(benchmark_beach())
        g.start()
        g.set_precision(prec)
        run_steps(g, 20)
        g.synchronize()
        c0 = g.poll(wait=True).copy()
        t0 = time.perf_counter()
        c, _ = run_steps(g, steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        c = g.poll(wait=True)
        out[name] = {"ms_per_step": wall / steps * 1e3, "sim_seconds": float(c[1] - c0[1]), "rejections": float(c[3] - c0[3]),
                     "pcg_iterations": float(c[18] - c0[18]), "failure": int(c[7])}
        g.close()
    out["identical_acceptance"] = (out["plain_fp32"]["sim_seconds"] == out["mixed_bf16_fp64"]["sim_seconds"] and
                                   out["plain_fp32"]["rejections"] == out["mixed_bf16_fp64"]["rejections"])
    return out


def refined_leg(steps: int = 40):
    """The refined beach 620x96x160 (9,523,200 cells): wall ms per accepted
    coupled step from the start, and the standalone pressure solve forms."""
    st = benchmark_beach(n=(620, 96, 160), spacing=(0.0125, 0.0125, 0.025))
    g = Solver(st)
    g.start()
    run_steps(g, 5)
    g.synchronize()
    t0 = time.perf_counter()
    c, _ = run_steps(g, steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    c = g.poll(wait=True)
    out = {"coupled_ms_per_step": wall / steps * 1e3, "steps": steps, "sim_time": float(c[1]), "failure": int(c[7])}
    g.close()
    out["pressure"] = pressure_leg(repeats=1, n=st.n, spacing=st.spacing, learned=False)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-pressure", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the in-loop mixed and refined-grid legs")
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

    # rooflines, timed live with HIP events on the solver's stream.  The dominant
    # kernel by step time is the coarse Chebyshev sweep family; its heaviest
    # launch is level 1.  Algorithmic bytes per launch (unique bytes at FP32):
    #   level-1 sweep: rows x (4 width coefficient bytes + x, b, D^-1, d, mask
    #     read + x', d written = 28 B);
    #   MacCormack forward trace: faces x (face velocity and valid flag read,
    #     limits, support, cached velocity and forward value written = 36 B);
    #   level-0 sweep: cells x 52 B (b, x, active, 6 coefficients, D^-1, d read;
    #     x', d written).
    N = s.cells
    nx, ny, nz = (int(v) for v in s.n)
    NF = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    lv = solver.levels()
    count1, width1 = lv[1][0] * lv[1][1] * lv[1][2], lv[1][11]
    copy_us = solver.time_kernel(3, reps=200)
    traffic_all = {}
    tfile = os.path.join(ROOT, "profiles", "traffic_kernels.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic_all = json.load(f)
    rooflines = []
    for key, kid, nbytes, label in (
            ("mg_coarse_cheb_level1", 4, count1 * (4 * width1 + 28),
             f"Chebyshev sweep on MG level 1 (k_mg_coarse<0>): {count1} rows x {width1}-point Galerkin stencil"),
            ("maccormack_forward", 5, NF * 36,
             f"MacCormack forward trace (op 27: k_maccormack_direct + queued k_maccormack_deferred): {NF} faces"),
            ("mg_cheb_level0", 0, 52 * N, f"level-0 Chebyshev sweep (k_mg_cheb): {N} cells")):
        us = solver.time_kernel(kid, reps=200)
        achieved = nbytes / (us * 1e-6) / 1e9
        rooflines.append({"bound": "hbm", "kernel": label, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_all.get(key, {}).get("hbm_bytes_per_launch"),
                          "avg_us": us, "bytes_per_launch": nbytes})
    rooflines[0]["copy_GBs"] = 8 * N / (copy_us * 1e-6) / 1e9
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
        "roofline": rooflines[0],
        "rooflines": rooflines,
    }
    solver.close()
    if not args.no_pressure:
        result["pressure"] = pressure_leg()
    if not args.no_extra:
        result["mixed_in_loop"] = inloop_leg()
        result["refined_620x96x160"] = refined_leg()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
