"""GPU: the submission scheduler at the boundary (water/work_scheduler.gd:41-87):
wv_advance with a batch budget and at most three submissions in flight, one
control readback per submission, GPU-time telemetry (gpu_ns, R_fluid), the
device-timed pressure sections per accepted step, and results independent of
how the batches are split into submissions (every op is gated on the device)."""
import numpy as np
import pytest

from waves_amd import solver as S
from waves_amd.settings import narrow_flume

pytestmark = pytest.mark.gpu


def test_budget_inflight_and_readbacks(gpu):
    s = narrow_flume(history_capacity=400)
    g = S.Solver(s)
    g.start()
    c = g.run_steps(100, max_batches=8)
    st = g.schedule_stats()
    print("after 100 steps:", {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}, flush=True)
    assert int(c[2]) == 100 and c[7] == 0
    # one readback per submission, fewer readbacks than accepted steps
    assert st["readbacks"] == st["submissions"] and st["inflight"] == 0
    assert st["readbacks"] <= 100
    assert 1 <= st["max_inflight"] <= S.MAX_INFLIGHT
    assert st["gpu_ms"] > 0 and st["ema_batch_ms"] > 0
    assert abs(st["r_fluid"] - st["gpu_ms"] * 1e-3 / float(c[1])) <= 1e-9 * st["r_fluid"] + 1e-12
    # the budget bounds a submission by the cost estimate (at least one batch)
    g.request(10_000)
    ema = st["ema_batch_ms"]
    n = g.frame(64, budget_ms=2.5 * ema)
    assert n == 2, (n, ema)
    assert g.frame(64, budget_ms=0.1 * ema) == 1
    # at most three unfinished submissions: a fourth is refused while the
    # first three (8 batches each) are still running
    g.poll(wait=True)
    got = [g.frame(8) for _ in range(4)]
    print("four submissions back to back:", got, flush=True)
    assert got[:3] == [8, 8, 8] and got[3] == 0
    g.poll(wait=True)
    st = g.schedule_stats()
    assert st["inflight"] == 0 and st["readbacks"] == st["submissions"]
    g.close()


def test_step_times_recorded(gpu):
    s = narrow_flume(history_capacity=200)
    g = S.Solver(s)
    g.start()
    g.run_steps(60, max_batches=8)
    t = g.step_times()
    assert t.shape == (60, 4)
    trial_ms, all_ms = t[:, 0] / 1e6, t[:, 1] / 1e6
    print(f"pressure ms per accepted step: median {np.median(trial_ms):.3f}, p90 {np.percentile(trial_ms, 90):.3f}, "
          f"max {trial_ms.max():.3f}; sections per step {np.unique(t[:, 3])}", flush=True)
    assert np.all(t[:, 0] > 0) and np.all(t[:, 1] >= t[:, 0]) and np.all(t[:, 3] >= 1)
    assert np.all(np.diff(t[:, 2].astype(np.int64)) >= 1)   # trial counts increase
    assert trial_ms.max() < 1000
    g.close()


def test_submission_split_does_not_change_results(gpu):
    """The same 40 steps advanced one batch per readback and eight batches
    per submission with three in flight end in bit-identical banks."""
    s = narrow_flume(history_capacity=100)
    a, b = S.Solver(s), S.Solver(s)
    a.start()
    b.start()
    for _ in range(40):
        a.step()
    b.run_steps(40, max_batches=8)
    ca, cb = a.control, b.control
    assert ca[2] == cb[2] == 40 and ca[1] == cb[1] and ca[3] == cb[3]
    for f in (S.FIELD_BANK_A, S.FIELD_BANK_B, S.FIELD_GEO):
        assert np.array_equal(a.read(f).view(np.uint32), b.read(f).view(np.uint32)), f
    a.close()
    b.close()
