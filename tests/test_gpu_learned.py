"""GPU: the learned warm start.

* What the trained network (waves_amd/data/warmstart_unet.npy, trained on the
  device by `python -m waves_amd.learned train`) gains over the same
  mixed-precision PCG without it, on pressure systems captured from a
  held-out narrow-flume run (amplitude 0.09 m, period 1.4 -> 1.5 s, neither in
  the training draws).  Both forms must reach the FP64 true relative residual
  1e-6; the learned start must save iterations on average.
* The capture -> train pipeline end to end on the device: systems captured
  from two short flume runs, a few hundred Adam steps from He-normal weights;
  the held-out loss must fall and the trained network must run inside the
  solver's mode 2 to the FP64 tolerance.
* The committed network on systems of the 310x48x80 beach: both forms must
  converge to 1e-6; the iterations and device time of each are reported (the
  network, trained on the flume, does not save iterations there)."""
import os

import numpy as np
import pytest

from waves_amd import learned as Lr

pytestmark = pytest.mark.gpu


def test_trained_network_saves_iterations(gpu):
    assert os.path.exists(Lr.WEIGHTS), "trained weights missing"
    params = np.load(Lr.WEIGHTS)
    st = Lr.run_settings(0.09, 1.5)
    res = Lr.evaluate(params, st, 320, 20, skip=20, tol=1e-6, modes=(1, 2), log=None)
    s = res["summary"]
    print(f"held-out flume: {s['systems']} systems; mixed {s['mode1']['mean_iterations']:.2f} it "
          f"{s['mode1']['mean_ms']:.2f} ms, mixed + learned {s['mode2']['mean_iterations']:.2f} it "
          f"{s['mode2']['mean_ms']:.2f} ms", flush=True)
    assert s["systems"] >= 10
    for m in (1, 2):
        assert s[f"mode{m}"]["converged"] == s["systems"]
        assert s[f"mode{m}"]["max_true_rel_residual"] <= 1.01e-6
    assert s["mode2"]["mean_iterations"] < s["mode1"]["mean_iterations"]


def test_capture_train_pipeline(gpu):
    n = (64, 48, 8)
    rng = np.random.default_rng(5)
    xs, ts = [], []
    for amp, per in ((0.06, 1.3), (0.09, 1.6)):
        st = Lr.run_settings(amp, per, n=n)
        recs = Lr.capture(st, 48, every=4, skip=8)
        assert len(recs) == 10
        for rec in recs:
            w = rec["wet"]
            # finite system; the target has unit wet RMS (or is zero)
            assert np.isfinite(rec["b"]).all() and np.isfinite(rec["r0"]).all()
            tg = Lr.target(rec["p_prev"], rec["p"], w)
            rms = np.sqrt((tg.astype(np.float64) ** 2 * w).sum() / w.sum())
            assert rms < 1e-12 or abs(rms - 1) < 1e-5
            xs.append(Lr.to_bf16_bits(Lr.features(rec["F"], rec["r0"], rec["p_prev"], w)).ravel())
            ts.append(tg)
    order = rng.permutation(len(ts))
    x, t = np.stack(xs)[order], np.stack(ts)[order]
    params, info = Lr.fit(x, t, n, batch=4, steps=300, lr=1e-3, val_fraction=0.2, log=lambda m: print(m, flush=True))
    assert np.isfinite(params).all()
    assert info["train_loss"] < 0.8 * info["val_loss_init"], info
    assert info["val_loss"] < info["val_loss_init"], info
    res = Lr.evaluate(params, Lr.run_settings(0.08, 1.5, n=n), 40, 10, skip=10, tol=1e-6, modes=(2,), log=None)
    s = res["summary"]
    assert s["systems"] >= 3 and s["mode2"]["converged"] == s["systems"]
    assert s["mode2"]["max_true_rel_residual"] <= 1.01e-6


def test_learned_start_on_beach(gpu):
    params = np.load(Lr.WEIGHTS)
    st = Lr.beach_settings(0.09, 1.5)
    res = Lr.evaluate(params, st, 1500, 150, skip=600, tol=1e-6, modes=(1, 2), log=None)
    s = res["summary"]
    print(f"310x48x80 beach: {s['systems']} systems; mixed {s['mode1']['mean_iterations']:.2f} it "
          f"{s['mode1']['mean_ms']:.2f} ms, mixed + learned {s['mode2']['mean_iterations']:.2f} it "
          f"{s['mode2']['mean_ms']:.2f} ms, saved {s['iterations_saved_vs_mixed']:+.2f} it", flush=True)
    assert s["systems"] >= 5
    for m in (1, 2):
        assert s[f"mode{m}"]["converged"] == s["systems"]
        assert s[f"mode{m}"]["max_true_rel_residual"] <= 1.01e-6
