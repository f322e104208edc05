#!/usr/bin/env python3
"""Benchmark: accepted simulated seconds per wall second of the coupled
free-surface step at 310x48x80 (the reference beach, water/lab_runner.gd:26,
plain FP32 method at the reference's own PCG tolerance 1e-4 and CFL 0.45) with
the seeded 1e-3 m alongshore perturbation.

A "step" is one accepted solver step (CFL substep trial accepted on the
device).  The solver is first run untimed from rest into the breaking phase
(t = 8 s: the waves shoal from about 4 s and break from about 6 s), then W
warmup steps, then exactly K accepted steps are timed, bracketed by device
synchronisation.  Submissions follow the reference's scheduler
(water/work_scheduler.gd): at most three in flight, one control readback each,
the quota stops the device after step K.  `value` is the timed window's
accepted sim-s / wall-s; `full_run` is the whole 20 sim-s run from rest, timed
in the same process, with its rate over the breaking phase (6 to 12 sim-s);
`window_scan` is the distribution of the same metric over consecutive windows
of K steps from 6 to 12 sim-s.  A step's simulated advance is set by the CFL
limit and some trials are rejected, so a short window varies several-fold
with where it starts (K = 20: p10 0.19, median 0.37, p90 0.61).

Also reported: metric 2, the in-loop pressure time per accepted step (device
clock around the pressure sections: hierarchy audit and rebuild, PCG,
projection, defect corrections), median and p90, for the window and the full
run; the standalone pressure solve at 310x48x80 at relative residual 1e-6 and
1e-8 (plain FP32 MGPCG, the mixed BF16/FP64 form, and the mixed form with the
learned warm start); the coupled step with the mixed PCG at 1e-6; the refined
620x96x160 beach; rooflines of the dominant kernels timed live with HIP
events on the solver's stream, beside a 1 GiB streaming copy; the CPU oracle
timed on the host's cores over a window of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from waves_amd.settings import benchmark_beach, default_beach  # noqa: E402
from waves_amd import solver as S  # noqa: E402
from waves_amd.solver import FIELD_QUEUES, Solver, stopped  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16
METRIC = "accepted sim-s per wall-s @310x48x80; pressure ms/step @1e-6 rel residual"
WINDOW_START = 8.0      # sim-s, breaking phase
FULL_RUN = 20.0         # sim-s


PROGRESS = False


def run_to_time(solver: Solver, t_end: float, max_batches: int = 8, curve=None):
    """Advance until the accepted time reaches t_end (request_time quota).
    curve: a list receiving (wall s, accepted sim-s) at each whole sim-second;
    with --progress each whole simulated second reached is printed to stderr."""
    solver.request_time(t_end)
    t0, mark = time.perf_counter(), 1.0
    while True:
        c = solver.poll(wait=False)
        st = solver.schedule_stats()
        c = solver._last_control() if c is None else c
        if c[1] >= mark:
            if curve is not None:
                curve.append((round(time.perf_counter() - t0, 3), round(float(c[1]), 3)))
            if PROGRESS:
                print(f"reached {float(c[1]):.3f} sim-s, {int(c[2])} steps, {time.perf_counter() - t0:.1f} s wall",
                      file=sys.stderr, flush=True)
            mark = float(np.floor(c[1])) + 1.0
        if c[1] >= t_end or stopped(c):
            if st["inflight"] == 0:
                return c
            solver.poll(wait=2)
            continue
        if st["inflight"] < S.MAX_INFLIGHT:
            solver.frame(max_batches)
        else:
            solver.poll(wait=2)


def pressure_ms(rows: np.ndarray):
    if len(rows) == 0:
        return None
    ms = rows[:, 0] / 1e6
    allms = rows[:, 1] / 1e6
    return {"median_ms": float(np.median(ms)), "p90_ms": float(np.percentile(ms, 90)), "mean_ms": float(ms.mean()),
            "with_rejected_trials_median_ms": float(np.median(allms)),
            "with_rejected_trials_p90_ms": float(np.percentile(allms, 90)), "steps": int(len(rows))}


def host_info():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count() or 1
    return nproc, model


def cpu_baseline(state, settings, steps: int, threads: int, model: str):
    """The CPU oracle (C restatement, OpenMP over `threads` host threads) on
    the benchmark window: the device's state at the window start loaded into
    the oracle (its multigrid rebuilt from the loaded operator), then `steps`
    accepted steps timed."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle.oracle import Oracle
    o = Oracle(settings)
    for f, v in state.items():
        o.set(f, v)
    t0 = time.perf_counter()
    c0 = o.control.copy()
    for _ in range(steps):
        o.step()
    wall = time.perf_counter() - t0
    c = o.control
    sim = float(c[1] - c0[1])
    o.close()
    return {"value": sim / wall, "unit": "accepted sim-s per wall-s", "cores": threads, "kind": "port",
            "cpu_model": model,
            "sample": f"CPU restatement (not lavapipe): oracle/ in C with OpenMP, {threads} threads on {model}; "
                      f"{steps} accepted steps from the window start t={float(c0[1]):.3f} s of the same 310x48x80 "
                      f"run ({sim:.5f} sim-s in {wall:.1f} s wall)"}


def capture_state(g: Solver):
    from oracle import oracle as O
    fields = (O.FIELD_BANK_A, O.FIELD_BANK_B, O.FIELD_WORK, O.FIELD_GEO, O.FIELD_CONTROL, O.FIELD_PLANE,
              O.FIELD_MOTION, O.FIELD_HISTORY, O.FIELD_COASTAL_HISTORY, O.FIELD_IMG_PLANES, O.FIELD_IMG_KINDS,
              O.FIELD_PROBE, O.FIELD_MG_VECTORS, O.FIELD_SNAPSHOT, O.FIELD_PREVIOUS)
    st = {f: g.read(f) for f in fields}
    audit = g.read(S.FIELD_AUDIT, np.uint32)
    audit[0] = 0      # the oracle keeps its coarse operator in its own layout: rebuild it
    st[O.FIELD_AUDIT] = audit
    return st


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
    if net is not None:
        out["net_layers"] = net_layer_roofline(net)
    g.close()
    if net is not None:
        net.close()
    return out


def net_layer_roofline(net, reps: int = 20):
    """Per-layer roofline of the warm-start network's inference (batch 1, on
    the features of the last learned solve): each 3x3x3 convolution is an
    implicit GEMM of 2 x cells x 27 cin x cout flops (bias excluded); its
    algorithmic bytes are the input read once (ca channels at the layer's
    level, cb at the next coarser level, bf16), the output written once (bf16
    after a ReLU, fp32 for the two heads) and the bf16 weights.  Times are HIP
    events around `reps` launches of the layer on the network's stream."""
    us = [float(v) for v in net.time_layers(reps)]
    rows, tot_f, tot_us = [], 0.0, 0.0
    names = ("e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "b3a", "b3b", "d2a", "d2b", "hc", "d1a", "d1b", "d0a", "d0b", "hf")
    for name, L, t in zip(names, net.layers(), us):
        cin = L["ca"] + L["cb"]
        flops = 2.0 * L["cells"] * 27 * cin * L["cout"]
        kp = (27 * cin + 1 + 15) // 16 * 16
        nbytes = (2.0 * (L["cells"] * L["ca"] + L["coarse_cells"] * L["cb"]) + L["cells"] * L["cout"] * (4 if L["cout"] == 1 else 2)
                  + 2.0 * L["cout"] * kp)
        tf = flops / (t * 1e-6) / 1e12
        gbs = nbytes / (t * 1e-6) / 1e9
        rows.append({"layer": name, "level": L["level"], "cells": L["cells"], "cin": cin, "cout": L["cout"], "us": float(t),
                     "tflops": tf, "frac_mfma_bf16": tf / MFMA_BF16_PEAK_TFLOPS, "gbs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS,
                     "intensity_flop_per_byte": flops / nbytes,
                     "bound": "mfma" if flops / nbytes > MFMA_BF16_PEAK_TFLOPS * 1e3 / HBM_PEAK_GBS else "hbm"})
        tot_f += flops
        tot_us += float(t)
    return {"layers": rows, "sum_us": tot_us, "gflop": tot_f / 1e9, "tflops": tot_f / (tot_us * 1e-6) / 1e12,
            "mfma": "v_mfma_f32_16x16x32_bf16 (K step 32) + one 16x16x16 step where kpad is 16 short of a multiple of 32"}


def inloop_leg(steps: int = 600, tol: float = 1e-6):
    """The coupled step at PCG tolerance `tol` with the batch program's PCG in
    plain FP32 and in the mixed form (wv_set_precision), same start: wall ms
    per accepted step, in-loop pressure ms per step and the acceptance record."""
    out = {"tolerance": tol, "steps": steps}
    for prec, name in ((0, "plain_fp32"), (1, "mixed_bf16_fp64")):
        g = Solver(benchmark_beach(pressure_tolerance=tol))
        g.start()
        g.set_precision(prec)
        g.run_steps(20)
        c0 = g.poll(wait=True).copy()
        t0 = time.perf_counter()
        c = g.run_steps(steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        c = g.poll(wait=True)
        rows = g.step_times()[int(c0[2]):]
        done = int(c[2] - c0[2])
        out[name] = {"ms_per_step": wall / max(done, 1) * 1e3, "steps": done, "sim_seconds": float(c[1] - c0[1]),
                     "rejections": float(c[3] - c0[3]), "pcg_iterations": float(c[18] - c0[18]),
                     "failure": int(c[7]), "pressure": pressure_ms(rows)}
        g.close()
    a, b = out["plain_fp32"], out["mixed_bf16_fp64"]
    out["identical_acceptance"] = a["sim_seconds"] == b["sim_seconds"] and a["rejections"] == b["rejections"]
    return out


def refined_leg(steps: int = 20):
    """The refined beach 620x96x160 (9,523,200 cells): wall ms per accepted
    coupled step from the start, and the standalone pressure solve forms."""
    st = benchmark_beach(n=(620, 96, 160), spacing=(0.0125, 0.0125, 0.025), history_capacity=256)
    g = Solver(st)
    g.start()
    g.run_steps(5)
    t0 = time.perf_counter()
    c = g.run_steps(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    out = {"coupled_ms_per_step": wall / steps * 1e3, "steps": steps, "sim_time": float(c[1]), "failure": int(c[7])}
    g.close()
    out["pressure"] = pressure_leg(repeats=1, n=st.n, spacing=st.spacing, learned=False)
    return out


def phase_rate(curve, t0: float, t1: float):
    """Accepted sim-s per wall-s between the first whole-second marks at or
    after t0 and t1 of a run's (wall s, sim-s) curve."""
    a = next((p for p in curve if p[1] >= t0), None)
    b = next((p for p in curve if p[1] >= t1), None)
    if a is None or b is None or b[0] <= a[0]:
        return None
    return (b[1] - a[1]) / (b[0] - a[0])


def scan_windows(steps: int, t0: float = 6.0, t1: float = 12.0):
    """Consecutive windows of `steps` accepted steps from t0 (after an
    untimed run from rest) until t1, each timed like the bench's window:
    the distribution of the windowed metric over the breaking phase."""
    s = benchmark_beach()
    g = Solver(s)
    g.start()
    run_to_time(g, t0)
    g.synchronize()
    vals = []
    while True:
        c = g.poll(wait=True)
        ts = float(c[1])
        w0 = time.perf_counter()
        c = g.run_steps(steps)
        g.synchronize()
        wall = time.perf_counter() - w0
        c = g.poll(wait=True)
        vals.append((float(c[1]) - ts) / wall)
        if PROGRESS and len(vals) % 25 == 0:
            print(f"{len(vals)} windows, t {float(c[1]):.3f} s", file=sys.stderr, flush=True)
        if float(c[1]) >= t1 or stopped(c):
            break
    g.close()
    v = np.array(vals)
    return {"steps_per_window": steps, "from_s": t0, "to_s": t1, "windows": len(vals), "median": float(np.median(v)),
            "p10": float(np.percentile(v, 10)), "p90": float(np.percentile(v, 90)), "min": float(v.min()),
            "max": float(v.max()), "mean": float(v.mean())}


def full_run(t_end: float = FULL_RUN, tol=None, precision: int = 0):
    """The whole 20 sim-s run from rest, timed (reset excluded); by default
    at the benchmark's PCG tolerance in plain FP32, else at `tol` and with
    the batch program's PCG in `precision` (1: BF16 V-cycle, FP64 dots)."""
    s = benchmark_beach() if tol is None else benchmark_beach(pressure_tolerance=tol)
    g = Solver(s)
    g.start()
    g.set_precision(precision)
    t0 = time.perf_counter()
    curve = []
    c = run_to_time(g, t_end, curve=curve)
    g.synchronize()
    wall = time.perf_counter() - t0
    c = g.poll(wait=True)
    st = g.schedule_stats()
    rows = g.step_times()
    out = {"value": float(c[1]) / wall, "unit": "accepted sim-s per wall-s", "sim_seconds": float(c[1]),
           "wall_s": wall, "steps": int(c[2]), "ms_per_step": wall / max(int(c[2]), 1) * 1e3,
           "rejections": int(c[3]), "pcg_iterations": int(c[18]), "failure": int(c[7]),
           "submissions": st["submissions"], "readbacks_per_accepted_step": st["readbacks"] / max(int(c[2]), 1),
           "gpu_s": st["gpu_ms"] / 1e3, "R_fluid": st["r_fluid"], "pressure": pressure_ms(rows),
           "wall_at_sim_second": curve, "breaking_phase_6_12s": phase_rate(curve, 6.0, 12.0),
           "pcg_tolerance": s.pressure_tolerance,
           "pcg_form": "mixed_bf16_fp64" if precision else "plain_fp32",
           "completed": bool(float(c[1]) >= t_end and int(c[7]) == 0)}
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--window-start", type=float, default=WINDOW_START)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-full", action="store_true", help="skip the full 20 sim-s run")
    ap.add_argument("--no-pressure", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the in-loop mixed and refined-grid legs")
    ap.add_argument("--save-state", help="write the state at the window start to this state file")
    ap.add_argument("--load-state", help="start the window from this state file instead of the untimed run")
    ap.add_argument("--progress", action="store_true", help="print each simulated second reached to stderr")
    ap.add_argument("--no-scan", action="store_true",
                    help="skip the distribution of --steps-step windows from 6 to 12 sim-s (run with the full run)")
    ap.add_argument("--full-tol", type=float, default=None, help="PCG tolerance of the full run (default: the benchmark's)")
    ap.add_argument("--full-precision", type=int, default=0, choices=(0, 1),
                    help="PCG of the full run's batch program: 0 plain FP32, 1 mixed (BF16 V-cycle, FP64 dots)")
    args = ap.parse_args()
    global PROGRESS
    PROGRESS = args.progress
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU (no multi-GPU path): --gpus must be 1"}))
        sys.exit(2)

    s = benchmark_beach()
    solver = Solver(s)
    solver.start()
    t_pre = time.perf_counter()
    if args.load_state:
        solver.load_state(args.load_state)
        c = solver.poll(wait=True)
    else:
        c = run_to_time(solver, args.window_start)
    preroll_wall = time.perf_counter() - t_pre
    if args.save_state:
        solver.save_state(args.save_state)
    cpu_state = None if args.no_cpu else capture_state(solver)
    c = solver.run_steps(args.warmup)
    solver.synchronize()
    c = solver.poll(wait=True)
    t_start_sim, steps_start, iters_start, rej_start = float(c[1]), int(c[2]), float(c[18]), float(c[3])
    st0 = solver.schedule_stats()
    t0 = time.perf_counter()
    c = solver.run_steps(args.steps)
    solver.synchronize()
    wall = time.perf_counter() - t0
    c = solver.poll(wait=True)
    st1 = solver.schedule_stats()
    sim = float(c[1]) - t_start_sim
    steps = int(c[2]) - steps_start
    value = sim / wall
    window_rows = solver.step_times()[steps_start:]

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
    copy_us = solver.time_kernel(6, reps=20)
    copy_gbs = 2 * (1 << 30) / (copy_us * 1e-6) / 1e9
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
                          "avg_us": us, "bytes_per_launch": nbytes, "frac_of_copy": achieved / copy_gbs})
        if kid == 5:  # faces the forward trace left to the queued general-path pass
            rooflines[-1]["queued_faces"] = int(solver.read(FIELD_QUEUES, np.uint32)[0])
    rooflines[0]["copy_GBs"] = copy_gbs
    rooflines[0]["copy"] = "streaming float4 copy, 1 GiB read + 1 GiB written"
    done_steps = max(steps, 1)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "accepted sim-s per wall-s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / done_steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: reference beach from rest, seeded 1e-3 m alongshore perturbation",
        "config": {"workload": f"coupled free-surface step, 310x48x80 beach (1,190,400 cells), plain FP32, PCG rel tol "
                               f"1e-4, CFL 0.45: {steps} accepted steps from t={t_start_sim:.3f} s (breaking phase, "
                               f"after an untimed run from rest to {args.window_start:g} s)",
                   "grid": [310, 48, 80], "sim_time_start": t_start_sim, "sim_seconds_timed": sim,
                   "preroll_wall_s": preroll_wall,
                   "pcg_iterations_per_step": (float(c[18]) - iters_start) / done_steps,
                   "rejections": float(c[3]) - rej_start, "failure": int(c[7]),
                   "submissions": st1["submissions"] - st0["submissions"],
                   "readbacks_per_accepted_step": (st1["readbacks"] - st0["readbacks"]) / done_steps,
                   "R_fluid_window": (st1["gpu_ms"] - st0["gpu_ms"]) * 1e-3 / max(sim, 1e-30)},
        "pressure_in_loop": pressure_ms(window_rows),
        "roofline": rooflines[0],
        "rooflines": rooflines,
    }
    solver.close()
    if not args.no_full:
        result["full_run"] = full_run(tol=args.full_tol, precision=args.full_precision)
    if not args.no_full and not args.no_scan:
        result["window_scan"] = scan_windows(args.steps)
    if not args.no_pressure:
        result["pressure"] = pressure_leg()
    if not args.no_extra:
        result["mixed_in_loop"] = inloop_leg()
        result["refined_620x96x160"] = refined_leg()
    if not args.no_cpu:
        nproc, model = host_info()
        result["cpu_baseline"] = cpu_baseline(cpu_state, s, args.cpu_steps, nproc, model)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
