"""GPU: the coupled step with the mixed-precision pressure solve in the batch
program (wv_set_precision 1: BF16 V-cycle operators, FP64 dot products)
against the reference's FP32 PCG (mode 0) on the narrow flume at PCG
tolerance 1e-6: two steps from the same state (a step's VOF transport uses
the previous step's velocities, so the second step is the first whose F sees
the changed pressure solve) stay within max |dF| <= 1e-4,
and over 100 steps the step acceptance (accepted steps, rejections, simulated
time) is identical."""
import numpy as np
import pytest

from waves_amd import solver as S
from waves_amd.settings import narrow_flume

pytestmark = pytest.mark.gpu


def advance(prec, steps, warm=20):
    st = narrow_flume()
    g = S.Solver(st)
    g.start()
    g.run_steps(warm)                 # same plain start for both
    g.synchronize()
    g.set_precision(prec)
    g.run_steps(steps)
    g.synchronize()
    c = g.poll(wait=True)
    F = g.read(int(c[0]))[:st.cells].copy()
    g.close()
    return c, F


def test_two_steps_match_plain(gpu):
    (c0, F0), (c1, F1) = advance(0, 2), advance(1, 2)
    d = float(np.abs(F0 - F1).max())
    print(f"two steps: max|dF| {d:.2e}, iterations plain {c0[18]:.0f} mixed {c1[18]:.0f}", flush=True)
    assert d <= 1e-4
    assert c0[2] == c1[2] and c0[1] == c1[1]


def test_acceptance_identical_over_100_steps(gpu):
    (c0, F0), (c1, F1) = advance(0, 100), advance(1, 100)
    print(f"100 steps: t {c0[1]:.5f} / {c1[1]:.5f}, rejections {c0[3]:.0f} / {c1[3]:.0f}, "
          f"iterations {c0[18]:.0f} / {c1[18]:.0f}, max|dF| {np.abs(F0 - F1).max():.2e}", flush=True)
    assert c0[7] == 0 and c1[7] == 0
    assert c0[2] == c1[2] and c0[3] == c1[3]
    assert abs(c0[1] - c1[1]) <= 1e-6 * c0[1]
