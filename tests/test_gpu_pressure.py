"""GPU: standalone pressure solve, plain FP32 MGPCG vs the mixed-precision
form (BF16 V-cycle operators, FP64 dots, FP64 true residual with iterative
refinement), against an FP64 sparse solve of the same system
(oracle/pressure64.py)."""
import numpy as np
import pytest

from waves_amd.settings import Settings
from waves_amd.scenarios import load_pressure_case
from waves_amd import solver as S
from oracle import pressure64

pytestmark = pytest.mark.gpu


def load_case(s, g):
    load_pressure_case(g, s)


def case_solver(n=(96, 48, 16)):
    s = Settings(n=n, origin=(0.5, -0.6, -0.4), zone_length=0.75, history_capacity=64)
    g = S.Solver(s)
    g.start()
    load_case(s, g)
    return s, g


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_plain_and_mixed_solves_match_fp64(gpu, tol):
    s, g = case_solver()
    res = {}
    for mode in (0, 1):
        load_case(s, g)
        r = g.pressure_solve(mode, tol)
        work, geo = g.read(S.FIELD_WORK), g.read(S.FIELD_GEO)
        A, b, wet = pressure64.assemble(work, geo, s.n)
        p = g.read(S.FIELD_PRESSURE64, np.float64)[wet]
        p64 = pressure64.solve(A, b)
        err = float(np.linalg.norm(p - p64) / np.linalg.norm(p64))
        relres = float(np.linalg.norm(b - A @ p) / np.linalg.norm(b))
        res[mode] = (r, err, relres)
        print(f"tol {tol:g} mode {mode}: iters {r['iterations']} restarts {r['restarts']} conv {r['converged']} "
              f"fail {r['failure']} true relres {r['true_rel_residual']:.3e} (cpu {relres:.3e}) "
              f"recurrence {r['recurrence_rel_residual']:.3e} |p-p64|/|p64| {err:.3e} device {r['device_ms']:.2f} ms",
              flush=True)
    for mode in (0, 1):
        r, err, relres = res[mode]
        assert r["failure"] == 0
        # the device's FP64 residual of its FP64 solution is the CPU's
        assert abs(r["true_rel_residual"] - relres) <= 1e-3 * relres + 1e-15
    # plain FP32: converged on the reference's recurrence criterion
    assert res[0][0]["converged"] == 1 and res[0][1] < 1e-3
    # mixed: converged on the FP64 true residual, down to 1e-8
    r, err, relres = res[1]
    assert r["converged"] == 1 and relres <= tol * 1.01
    assert err < 1e3 * tol
    g.close()
