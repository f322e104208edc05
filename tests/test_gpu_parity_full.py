"""GPU parity at the production grid, 310x48x80 (1,190,400 cells): the
reference beach run on the device into the breaking phase, its state loaded
into the oracle and the device alike (the identical-input harness), then

* one batch op group by op group, each group held to the parity bar (rational
  groups bit-exact, transcendental ones within 2e-5) -- this is where the
  planner picks the large-grid kernels: one lane per coarse row on level 1,
  one lane per fine row in Galerkin stage 1, the one-workgroup tail on levels
  4-6;
* one full accepted step run independently on both sides: cell and solid
  flags and exported kinds bit-exact, max |dF| <= 1e-4, identical acceptance.

Both sides rebuild the multigrid hierarchy from the loaded fine operator (the
audit state's "hierarchy exists" word cleared), since the oracle keeps its
coarse operator in the reference's 125-slot layout."""
import sys
import time

import numpy as np
import pytest

from waves_amd.settings import benchmark_beach
from waves_amd import solver as S
from tests import parity as P
from tests.test_gpu_parity import mg_report

pytestmark = pytest.mark.gpu

T_START = 6.0   # sim-s: shoaling into breaking (the waves reach the toe at about 4 s)
LOAD = (P.O.FIELD_BANK_A, P.O.FIELD_BANK_B, P.O.FIELD_WORK, P.O.FIELD_GEO, P.O.FIELD_CONTROL, P.O.FIELD_PLANE,
        P.O.FIELD_MOTION, P.O.FIELD_HISTORY, P.O.FIELD_COASTAL_HISTORY, P.O.FIELD_IMG_PLANES, P.O.FIELD_IMG_KINDS,
        P.O.FIELD_PROBE, P.O.FIELD_MG_VECTORS, P.O.FIELD_SNAPSHOT, P.O.FIELD_PREVIOUS)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def breaking_state(gpu):
    s = benchmark_beach()
    g = S.Solver(s)
    g.start()
    g.request_time(T_START)
    t0 = time.time()
    while True:
        g.frame(4)
        c = g.poll(wait=True)
        if c[1] >= T_START or S.stopped(c):
            break
    g.synchronize()
    c = g.control
    say(f"pre-roll to t = {c[1]:.3f} s: {int(c[2])} accepted steps, {int(c[3])} rejections, "
        f"{time.time() - t0:.1f} s wall, trial open {c[4]}")
    assert c[7] == 0 and c[4] == 0
    state = {f: g.read(f) for f in LOAD}
    audit = g.read(S.FIELD_AUDIT, np.uint32)
    audit[0] = 0          # no hierarchy yet: the next build is a full one on both sides
    state[P.O.FIELD_AUDIT] = audit
    g.close()
    return s, state


def loaded_pair(s, state):
    pair = P.Pair(s, start=False)
    pair.load(state)
    return pair


@pytest.mark.timeout(600)
def test_full_grid_batch_groups(breaking_state):
    s, state = breaking_state
    pair = loaded_pair(s, state)
    o, g = pair.o, pair.g
    assert g.levels() == o.levels()
    lv = g.levels()
    say("levels", [m[8] for m in lv])
    quota = float(o.control[2]) + 1
    o.set_quota(quota)
    g.request(int(quota))
    for k, grp in enumerate(P.batch_groups(8)):
        t = time.time()
        rep = pair.group(grp)
        if "build" in grp:
            mg_report(o, g)
        bad = {name: v for name, v in rep.items() if v[0]}
        say(k, "group", grp, f"{time.time() - t:.2f} s", "c[4..8]", o.control[4:9], bad if bad else "exact")
        P.check_group(grp, rep, "310x48x80")
    c = o.control
    say("after one batch: accepted", c[2], "rejections", c[3], "failure", c[7], "iterations", c[8], "defects", c[58])
    pair.close()


@pytest.mark.timeout(600)
def test_full_grid_one_step(breaking_state):
    s, state = breaking_state
    pair = loaded_pair(s, state)
    o, g = pair.o, pair.g
    t = time.time()
    bo = o.step()
    say(f"oracle step: {bo} batches, {time.time() - t:.1f} s")
    bg = g.step()
    co, cg = o.control, g.control
    say("batches", bo, bg, "accepted", co[2], cg[2], "rejections", co[3], cg[3], "retries", co[29], cg[29],
        "failure", co[7], cg[7], "PCG iterations", co[18], cg[18], "defects", co[58], cg[58])
    assert co[2] == cg[2] and co[3] == cg[3] and co[7] == cg[7] == 0
    assert co[0] == cg[0]
    acc = int(co[0])
    Fo, Fg = o.bank(acc)[: o.N], g.read(acc)[: o.N]
    dF = float(np.abs(Fo - Fg).max())
    ko, kg = o.get(P.O.FIELD_GEO)[1::2], g.read(S.FIELD_GEO)[1::2]
    io, ig = o.get(P.O.FIELD_IMG_KINDS), g.read(S.FIELD_IMG_KINDS)
    say("max|dF|", dF, "kind flags differing", int((ko != kg).sum()), "exported kinds differing", int((io != ig).sum()))
    assert dF <= 1e-4
    assert np.array_equal(ko, kg) and np.array_equal(io, ig)
    # liquid volume against the reference's ledger (pressure_pass.glslinc:91-95)
    cell = float(np.prod(s.spacing))
    for c, F in ((co, Fo), (cg, Fg)):
        V = float(F.astype(np.float64).sum()) * cell
        assert abs(c[28]) <= 1e-3 * c[26]
        say("volume", V, "c[17]", c[17], "mass defect c[28]", c[28])
    pair.close()
