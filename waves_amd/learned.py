"""Training data, training and evaluation of the learned warm start.

No reference counterpart.  A pressure system of the coupled step is captured
at an accepted step: the accepted state before the step (volume fraction F and
pressure p_prev), the right-hand side b and the face coefficients the accepted
trial assembled (work fields 5..12), the wet mask and the accepted pressure
p*.  The network sees F, r0 = b - A p_prev, p_prev and the mask (each scaled by
its wet RMS) and is trained to predict the shape of the correction
p* - p_prev; the solver scales it by a Galerkin projection, so only the shape
matters (waves_amd/csrc/wv_solve.hip).

Usage on a GPU host:
    python -m waves_amd.learned train --out waves_amd/data/warmstart_unet.npy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Dict, List

import numpy as np

from . import solver as S
from .net import Net, to_bf16_bits
from .settings import Settings, narrow_flume, default_beach

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "warmstart_unet.npy")


def masked_operator(work: np.ndarray, wet: np.ndarray, p: np.ndarray, n) -> np.ndarray:
    """A p in FP64 for the masked 7-point operator of pressure_core.glslinc:120-125:
    sum over the six faces of coefficient * (p_here - p_neighbour), neighbours
    outside x/y or dry contribute 0, z is periodic.  Zero on dry cells."""
    nx, ny, nz = (int(v) for v in n)
    N = nx * ny * nz
    W = np.asarray(work, np.float64).reshape(14, N)
    w = wet.reshape(nz, ny, nx)
    pw = np.where(w, np.asarray(p, np.float64).reshape(nz, ny, nx), 0.0)
    out = np.zeros_like(pw)
    face = 0
    for ax in (2, 1, 0):                       # x, y, z in array order (nz, ny, nx)
        for sign in (-1, 1):
            if ax == 0:
                nb = np.roll(pw, -sign, axis=0)
            else:
                nb = np.zeros_like(pw)
                src = [slice(None)] * 3
                dst = [slice(None)] * 3
                if sign < 0:
                    dst[ax], src[ax] = slice(1, None), slice(None, -1)
                else:
                    dst[ax], src[ax] = slice(None, -1), slice(1, None)
                nb[tuple(dst)] = pw[tuple(src)]
            out += W[6 + face].reshape(nz, ny, nx) * (pw - nb)
            face += 1
    return np.where(w, out, 0.0).ravel()


def features(F: np.ndarray, r: np.ndarray, p: np.ndarray, wet: np.ndarray) -> np.ndarray:
    """(cells, 8) float32 network input: F, r / rms_wet(r), p / rms_wet(p), wet, 0 x 4."""
    w = wet.astype(np.float64)
    cnt = max(w.sum(), 1.0)
    sr = np.sqrt((np.asarray(r, np.float64) ** 2 * w).sum() / cnt)
    sp = np.sqrt((np.asarray(p, np.float64) ** 2 * w).sum() / cnt)
    x = np.zeros((len(F), 8), np.float32)
    x[:, 0] = F
    if sr > 0:
        x[:, 1] = np.asarray(r, np.float64) / sr * w
    if sp > 0:
        x[:, 2] = np.asarray(p, np.float64) / sp * w
    x[:, 3] = w
    return x


def target(p_prev: np.ndarray, p_star: np.ndarray, wet: np.ndarray) -> np.ndarray:
    d = (np.asarray(p_star, np.float64) - p_prev) * wet
    s = np.sqrt((d ** 2).sum() / max(wet.sum(), 1))
    return (d / s if s > 0 else d).astype(np.float32)


def capture(settings: Settings, steps: int, every: int = 1, skip: int = 0, keep_work: bool = False,
            log=None) -> List[Dict]:
    """Run `steps` accepted steps of the coupled solver and capture the main
    pressure system of every `every`-th step after the first `skip`.

    A step's trial solves twice: the main solve from the previous pressure,
    then (when the post-projection check asks for it, coupled_motion.glslinc
    ops 63-65) a defect correction from zero whose right-hand side replaces b
    in work field 5; the main solve's pressure is then kept in work field 13
    and added back.  The main system is rebuilt as (A, b1 = A p_first) with
    p_first the main solve's converged pressure, so p_first solves it exactly
    and the step's own residual (at most the tolerance) is not carried over."""
    g = S.Solver(settings)
    g.start()
    N = settings.cells
    out = []
    try:
        for k in range(steps):
            c = g.control
            if S.stopped(c):
                break
            take = k >= skip and (k - skip) % every == 0
            if take:
                p_prev = g.read(int(c[0]))[-N:].copy()
            g.step()
            if not take:
                continue
            c = g.control
            bank = g.read(int(c[0]))
            work = g.read(S.FIELD_WORK)
            geo = g.read(S.FIELD_GEO)
            wet = geo.reshape(N, 2)[:, 1] > 0.5
            saved = work[13 * N:14 * N]
            defect = bool(np.any(saved != 0))
            p_first = np.where(wet, saved if defect else bank[-N:], 0).astype(np.float32)
            b1 = masked_operator(work, wet, p_first, settings.n).astype(np.float32)
            p0 = np.where(wet, p_prev, 0).astype(np.float32)
            r0 = b1 - masked_operator(work, wet, p0, settings.n)
            rec = {"F": bank[:N].copy(), "p_prev": p0, "p": p_first, "b": b1, "wet": wet, "defect": defect,
                   "r0": r0.astype(np.float32), "start_relres": float(np.linalg.norm(r0) / max(np.linalg.norm(b1), 1e-30)),
                   "step": int(c[2]), "t": float(c[1])}
            if keep_work:
                w = work.copy()
                w[5 * N:6 * N] = b1
                rec["work"] = w
                rec["geo"] = geo.copy()
            out.append(rec)
            if log and len(out) % 50 == 0:
                log(f"captured {len(out)} systems (step {int(c[2])}, t {float(c[1]):.2f} s)")
    finally:
        g.close()
    return out


def run_settings(amplitude: float, period: float, n=None, tol: float = 1e-6) -> Settings:
    base = narrow_flume(amplitude=amplitude, period=period, pressure_tolerance=tol)
    if n is not None:
        base.n = n
    return base


def beach_settings(amplitude: float, period: float, tol: float = 1e-4) -> Settings:
    """The 310x48x80 beach of the benchmark (seeded 1e-3 m perturbation) with
    another incident wave; PCG tolerance as in the benchmark's loop."""
    return default_beach(amplitude=amplitude, period=period, pressure_tolerance=tol, perturbation=1e-3, seed=1)


def build_dataset(runs: int = 8, steps: int = 1000, every: int = 4, skip: int = 0, seed: int = 2024, log=print,
                  scenario: str = "flume"):
    """(x bf16 bits (P, cells*8), t (P, cells), meta) from `runs` runs with
    seeded amplitudes in [0.04, 0.10] m and periods in [1.2, 1.8] s, on the
    narrow flume or (scenario "beach") the 310x48x80 beach."""
    rng = np.random.default_rng(seed)
    xs, ts, meta = [], [], []
    for r in range(runs):
        amp, per = float(rng.uniform(0.04, 0.10)), float(rng.uniform(1.2, 1.8))
        st = beach_settings(amp, per) if scenario == "beach" else run_settings(amp, per)
        t0 = time.time()
        recs = capture(st, steps, every, skip)
        for rec in recs:
            xs.append(to_bf16_bits(features(rec["F"], rec["r0"], rec["p_prev"], rec["wet"])).ravel())
            ts.append(target(rec["p_prev"], rec["p"], rec["wet"]))
        rel = [rec["start_relres"] for rec in recs]
        meta.append({"amplitude": amp, "period": per, "pairs": len(recs),
                     "median_start_relres": float(np.median(rel)) if rel else 0.0,
                     "defect_fraction": float(np.mean([rec["defect"] for rec in recs])) if recs else 0.0})
        log(f"run {r}: amplitude {amp:.3f} m period {per:.2f} s -> {len(recs)} pairs in {time.time() - t0:.1f} s, "
            f"median |b - A p_prev| / |b| {meta[-1]['median_start_relres']:.2e}, defect solves {meta[-1]['defect_fraction']:.2f}")
    return np.stack(xs), np.stack(ts), meta


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero bias and padding, in the layout of include/waves_net.h."""
    rng = np.random.default_rng(seed)
    shapes = [(8, 8), (8, 8), (16, 8), (16, 16), (32, 16), (32, 32), (64, 32), (64, 64),
              (32, 96), (32, 32), (1, 32), (16, 48), (16, 16), (8, 24), (8, 8), (1, 8)]
    parts = []
    for co, ci in shapes:
        kp = (27 * ci + 1 + 15) // 16 * 16
        w = np.zeros((co, kp), np.float32)
        w[:, :27 * ci] = rng.normal(0, np.sqrt(2.0 / (27 * ci)), (co, 27 * ci))
        parts.append(w.ravel())
    return np.concatenate(parts)


def fit(x: np.ndarray, t: np.ndarray, n, batch: int = 8, steps: int = 4000, lr: float = 1e-3,
        val_fraction: float = 0.1, seed: int = 7, init_seed: int = 0, log=print, init=None):
    """Adam on the device; the last `val_fraction` of the pairs is held out.
    Learning rate: `lr` for the first 70 % of the steps, then lr / 10.  The
    weights start from `init` (a parameter vector, to fine-tune) or from
    He-normal draws."""
    P = len(t)
    n_val = max(1, int(P * val_fraction))
    n_train = P - n_val
    net = Net(n, batch)
    net.set_params(init_params(init_seed) if init is None else np.asarray(init, np.float32))
    net.set_dataset(x, t)
    v0 = net.eval(n_train, n_val)
    s1 = int(steps * 0.7)
    l1 = net.train(s1, n_train, lr, seed)
    l2 = net.train(steps - s1, n_train, lr / 10, seed + 1)
    v1 = net.eval(n_train, n_val)
    tr = net.eval(0, n_train)
    log(f"training: {steps} steps, batch {batch}, {n_train} train / {n_val} held-out pairs; "
        f"held-out loss {v0:.4f} -> {v1:.4f}, train loss {tr:.4f}")
    params = net.get_params()
    net.close()
    return params, {"steps": steps, "batch": batch, "train_pairs": n_train, "val_pairs": n_val,
                    "val_loss_init": v0, "val_loss": v1, "train_loss": tr,
                    "loss_curve": np.concatenate([l1, l2])[:: max(1, steps // 100)].tolist()}


def replay(g: S.Solver, s: Settings, rec: Dict, mode: int, tol: float) -> Dict:
    """Solve one captured system on solver `g` (same grid) from p_prev."""
    N = s.cells
    bank = np.zeros(g.field_bytes(S.FIELD_BANK_A) // 4, np.float32)
    bank[:N] = rec["F"]
    bank[-N:] = rec["p_prev"]
    g.write(S.FIELD_BANK_A, bank)
    g.write(S.FIELD_BANK_B, bank)
    g.write(S.FIELD_GEO, rec["geo"])
    g.write(S.FIELD_WORK, rec["work"])
    g.synchronize()
    return g.pressure_solve(mode, tol, given_rhs=True)


def evaluate(params: np.ndarray, settings: Settings, steps: int, every: int, skip: int = 0, tol: float = 1e-6,
             modes=(0, 1, 2), log=print) -> Dict:
    """Capture held-out systems from a run of `settings` and solve each from
    p_prev in every mode (one untimed solve per system first, so the
    hierarchy for its geometry is built outside the timed solve)."""
    held = capture(settings, steps, every, skip, keep_work=True, log=log)
    g = S.Solver(settings)
    g.start()
    net = Net(settings.n, 1)
    net.set_params(params)
    g.attach_net(net)
    rows = []
    try:
        for rec in held:
            row = {"step": rec["step"], "t": rec["t"], "start_relres": rec["start_relres"]}
            for mode in modes:
                replay(g, settings, rec, mode, tol)
                r = replay(g, settings, rec, mode, tol)
                row[f"mode{mode}"] = {k: r[k] for k in ("iterations", "restarts", "converged", "device_ms", "true_rel_residual")}
            rows.append(row)
            rec.pop("work", None)
    finally:
        g.close()
        net.close()
    summary = {"systems": len(rows), "tolerance": tol}
    for mode in modes:
        it = [row[f"mode{mode}"]["iterations"] for row in rows]
        ms = [row[f"mode{mode}"]["device_ms"] for row in rows]
        rr = [row[f"mode{mode}"]["true_rel_residual"] for row in rows]
        summary[f"mode{mode}"] = {"mean_iterations": float(np.mean(it)), "mean_ms": float(np.mean(ms)),
                                  "median_true_rel_residual": float(np.median(rr)), "max_true_rel_residual": float(np.max(rr)),
                                  "converged": int(sum(row[f"mode{mode}"]["converged"] for row in rows))}
    if 0 in modes and 2 in modes:
        summary["speedup_plain_over_learned"] = summary["mode0"]["mean_ms"] / summary["mode2"]["mean_ms"]
    if 1 in modes and 2 in modes:
        summary["speedup_mixed_over_learned"] = summary["mode1"]["mean_ms"] / summary["mode2"]["mean_ms"]
        summary["iterations_saved_vs_mixed"] = summary["mode1"]["mean_iterations"] - summary["mode2"]["mean_iterations"]
    return {"summary": summary, "systems": rows}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", default=WEIGHTS)
    tr.add_argument("--runs", type=int, default=8)
    tr.add_argument("--steps", type=int, default=1000)
    tr.add_argument("--every", type=int, default=4)
    tr.add_argument("--skip", type=int, default=0, help="accepted steps of each run before the first capture")
    tr.add_argument("--train-steps", type=int, default=4000)
    tr.add_argument("--batch", type=int, default=8)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--scenario", choices=("flume", "beach"), default="flume")
    tr.add_argument("--init", default=None, help="weights to fine-tune from")
    ev = sub.add_parser("eval")
    ev.add_argument("--weights", default=WEIGHTS)
    ev.add_argument("--out", required=True)
    ev.add_argument("--amplitude", type=float, default=0.09)
    ev.add_argument("--period", type=float, default=1.5)
    ev.add_argument("--sim-seconds", type=float, default=10.0)
    ev.add_argument("--every", type=int, default=40)
    ev.add_argument("--tol", type=float, default=1e-6)
    args = ap.parse_args(argv)
    if args.cmd == "eval":
        params = np.load(args.weights)
        st = default_beach(amplitude=args.amplitude, period=args.period, pressure_tolerance=1e-4,
                           perturbation=1e-3, seed=1)
        # steps covering the requested simulated time at the CFL-limited step (about 1/300 s)
        steps = int(args.sim_seconds * 300)
        t0 = time.time()
        res = evaluate(params, st, steps, args.every, skip=args.every, tol=args.tol,
                       log=lambda m: print(m, flush=True))
        res["summary"].update({"grid": list(st.n), "amplitude": args.amplitude, "period": args.period,
                               "in_loop_tolerance": 1e-4, "seconds": time.time() - t0,
                               "sim_time_last": res["systems"][-1]["t"] if res["systems"] else 0})
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res["summary"]))
        return 0
    if args.cmd == "train":
        t0 = time.time()
        x, t, meta = build_dataset(args.runs, args.steps, args.every, skip=args.skip, scenario=args.scenario)
        n = (beach_settings(0.08, 1.4) if args.scenario == "beach" else run_settings(0.08, 1.4)).n
        init = np.load(args.init) if args.init else None
        params, info = fit(x, t, n, args.batch, args.train_steps, args.lr, init=init)
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        np.save(args.out, params)
        info.update({"runs": meta, "pairs": int(len(t)), "grid": list(n), "scenario": args.scenario,
                     "init": os.path.basename(args.init) if args.init else None, "seconds": time.time() - t0})
        with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
            json.dump(info, f, indent=1)
        print(json.dumps({k: v for k, v in info.items() if k != "loss_curve"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
