"""Summation order alone, on the oracle (test infrastructure, run by hand:
`python tests/oracle_scatter_order.py [steps]`).

The reference scatters the Galerkin coarse operator with float
compare-and-swap atomics (pressure_multigrid.glslinc:128-134): the order of the
additions into a coarse entry is whatever the GPU's scheduling makes it, so two
runs of the reference are not bitwise reproducible.  This script runs the
oracle twice on the narrow flume (124x48x8, plain FP32, PCG 1e-6, the
configuration of tests/test_gpu_configs.py::test_narrow_flume_300_steps_against_oracle),
once with the scatter's additions in ascending and once in descending entry
order (wo_set_scatter_order), everything else identical, and prints the same
per-50-step comparison as that GPU test: accepted time, rejections, PCG
iterations, defect corrections and max |dF| between the two trajectories.
The divergence it shows is what two valid executions of the reference itself
differ by, the yardstick for the device-vs-oracle drift over independent
trajectories."""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle.oracle import Oracle  # noqa: E402
from waves_amd.settings import narrow_flume  # noqa: E402


def main(steps: int = 300) -> int:
    s = narrow_flume()
    a, b = Oracle(s), Oracle(s)
    b.set_scatter_order(True)
    a.initial()
    b.initial()
    t0, worst = time.time(), 0.0
    for k in range(1, steps + 1):
        a.step()
        b.step()
        ca, cb = a.control, b.control
        if ca[7] != 0 or cb[7] != 0:
            print(f"step {k}: failure codes {ca[7]:.0f} / {cb[7]:.0f}")
            return 1
        if k % 50 == 0 or ca[3] != cb[3]:
            Fa, Fb = a.bank(int(ca[0]))[: a.N], b.bank(int(cb[0]))[: b.N]
            dF = float(np.abs(Fa - Fb).max())
            worst = max(worst, dF)
            print(f"step {k}: t {ca[1]:.5f} / {cb[1]:.5f}, rejections {ca[3]:.0f} / {cb[3]:.0f}, PCG iterations "
                  f"{ca[18]:.0f} / {cb[18]:.0f}, defects {ca[58]:.0f} / {cb[58]:.0f}, max|dF| {dF:.3g} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    print(f"largest max|dF| over the run: {worst:.3g}")
    print("exported kinds differing:", int((a.get(8) != b.get(8)).sum()))
    a.close()
    b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(int(sys.argv[1]) if len(sys.argv) > 1 else 300))
