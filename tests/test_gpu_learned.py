"""GPU: what the trained warm-start network (waves_amd/data/warmstart_unet.npy,
trained on the device by `python -m waves_amd.learned train`) gains over the
same mixed-precision PCG without it, on pressure systems captured from a
held-out narrow-flume run (amplitude 0.09 m, period 1.4 -> 1.5 s, neither in
the training draws).  Both forms must reach the FP64 true relative residual
1e-6; the learned start must save iterations on average."""
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
