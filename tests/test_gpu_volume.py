"""GPU: the FP64 global liquid volume (wv_liquid_volume) closes the
reference's volume ledger (pressure_pass.glslinc:91-95) to a relative drift
<= 1e-6 per accepted step over 60 steps of the narrow flume with waves, and
agrees with the reference's FP32 volume c[17]."""
import numpy as np
import pytest

from waves_amd import solver as S
from waves_amd.settings import narrow_flume

pytestmark = pytest.mark.gpu


def defect64(V, c):
    # c[28] after acceptance: V + outflow - V0 - roundoff + relaxation terms
    c = c.astype(np.float64)
    return V + c[40] - c[26] - c[24] + c[74] + c[66] - c[73] - c[65] - c[72] - c[64]


def test_volume_ledger_closes_in_fp64(gpu):
    st = narrow_flume(pressure_tolerance=1e-4)
    g = S.Solver(st)
    g.start()
    rows = []
    for k in range(60):
        g.step()
        c = g.control
        assert not S.stopped(c)
        V = g.liquid_volume(int(c[0]))
        rows.append((V, defect64(V, c), float(c[17])))
    g.close()
    V = np.array([r[0] for r in rows])
    d = np.array([r[1] for r in rows])
    drift = np.abs(np.diff(d)) / V[1:]
    print(f"volume {V[0]:.6f} .. {V[-1]:.6f} m^3, max per-step relative drift after ledgers {drift.max():.2e}, "
          f"FP32 c[17] vs FP64 rel {np.max(np.abs(np.array([r[2] for r in rows]) - V) / V):.2e}", flush=True)
    assert drift.max() <= 1e-6
    assert np.max(np.abs(np.array([r[2] for r in rows]) - V) / V) <= 1e-5
