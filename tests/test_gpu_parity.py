"""GPU parity: every op group of a batch on identical inputs, and short
trajectories, HIP solver vs CPU oracle (water/shaders/compute/*.glslinc
semantics).  Rational groups are bit-exact; groups that evaluate
transcendentals agree to 2e-5 relative (tests/parity.py: TRANSCENDENTAL)."""
import sys

import numpy as np
import pytest

from waves_amd.settings import narrow_flume
from tests import parity as P

pytestmark = pytest.mark.gpu

# 64 x 48 x 4 flume: beach toe inside, ocean zone, a 2-cell periodic axis on MG level 2
SMALL = dict(n=(64, 48, 4), history_capacity=64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_initial_commands_match(gpu):
    s = narrow_flume(**SMALL)
    o, g = P.make_pair(s)
    P.start_pair(o, g)
    rep = P.compare_named(o, g)
    for k, v in rep.items():
        say("initial", k, v)
    # rational ops are bit-exact; the probes evaluate sinh/cosh/tanh/cos/sin
    for name in ("bankA", "bankB", "work", "geo", "plane", "motion", "img_planes", "img_kinds", "control"):
        assert rep[name][0] == 0, (name, rep[name])
    assert rep["probe"][2] < 1e-5
    assert g.levels() == o.levels()
    g.close()
    o.close()


def mg_report(o, g, tol=1e-5):
    """After a build: active masks bit-exact, l1 scales and coarse rows within
    FP32 reordering of the Galerkin sums."""
    vo, vg = o.get(P.O.FIELD_MG_VECTORS), g.read(P.S.FIELD_MG_VECTORS)
    ro = o.get(P.O.FIELD_MG_ROWS)[0::2]
    rg = g.read(P.S.FIELD_MG_ROWS)
    for l, m in enumerate(o.levels()):
        off, cnt = m[3], m[8]
        act_o, act_g = vo[8 * off + 3 * cnt: 8 * off + 4 * cnt], vg[8 * off + 3 * cnt: 8 * off + 4 * cnt]
        assert np.array_equal(act_o, act_g), ("active mask", l)
        d_o, d_g = vo[8 * off + 4 * cnt: 8 * off + 5 * cnt], vg[8 * off + 4 * cnt: 8 * off + 5 * cnt]
        rel = float(np.abs(d_o - d_g).max() / max(np.abs(d_o).max(), 1e-30))
        say("  level", l, "count", cnt, "l1 rel diff", rel)
        assert rel < tol if l > 0 else rel == 0
    rel = float(np.abs(ro - rg).max() / max(np.abs(ro).max(), 1e-30)) if ro.size else 0.0
    say("  coarse rows rel diff", rel)
    assert rel < tol
    # the coarse sweeps address every row by delta + one periodic correction,
    # which needs the entries past a non-periodic edge to be exactly zero
    for r0, N, W, mask in P.coarse_edge_entries(o):
        assert np.all(rg[r0: r0 + W * N].reshape(W, N)[mask] == 0), ("edge entries", r0)


def run_batch(pair, quota=1.0, groups=None, mutate=None, where="", mg_tol=1e-5, transcendental_all=False):
    """One batch group by group on identical inputs, every group held to the
    parity bar.  mutate: {group index: fn(c)} forced control states."""
    o, g = pair.o, pair.g
    o.set_quota(quota)
    g.request(int(quota))
    for k, grp in enumerate(groups or P.batch_groups()):
        rep = pair.group(grp, (mutate or {}).get(k))
        if "build" in grp:
            mg_report(o, g, mg_tol)
        bad = {name: v for name, v in rep.items() if v[0]}
        say(where, k, "group", grp, "c[4..8]", o.control[4:9], bad if bad else "exact")
        P.check_group(grp, rep, where, transcendental_all)
        if transcendental_all:       # cell flags stay bit-exact
            assert np.array_equal(o.get(P.O.FIELD_GEO)[1::2], g.read(P.S.FIELD_GEO)[1::2]), (where, grp)
    return o.control


def test_batch_groups_on_identical_inputs(gpu):
    pair = P.Pair(narrow_flume(**SMALL))
    assert pair.g.levels() == pair.o.levels()
    run_batch(pair, where="default")
    pair.close()


def test_batch_groups_one_lane_per_row_no_tail(gpu, monkeypatch):
    """The code paths the planner picks for the large grid, forced on the small
    one: one lane per coarse row (the k_mg_coarse G == 1 branch), one lane per
    fine row in Galerkin stage 1 (ga = 1), every coarse level launched on its
    own instead of the one-workgroup tail."""
    monkeypatch.setenv("WAVES_MG_LANE_TARGET", "1")
    monkeypatch.setenv("WAVES_MG_TAIL", "0")
    pair = P.Pair(narrow_flume(**SMALL))
    monkeypatch.delenv("WAVES_MG_LANE_TARGET")
    monkeypatch.delenv("WAVES_MG_TAIL")
    run_batch(pair, where="lanes1")
    pair.close()


def test_short_trajectory(gpu):
    s = narrow_flume(**SMALL)
    o, g = P.make_pair(s)
    P.start_pair(o, g)
    for k in range(3):
        bo = o.step()
        bg = g.step()
        co, cg = o.control, g.control
        say("step", k, "batches", bo, bg, "accepted", co[2], cg[2], "t", co[1], cg[1], "iters", co[18], cg[18],
            "fail", co[7], cg[7], "div", co[14], cg[14])
        assert co[2] == cg[2] and cg[7] == 0
        Fo = o.bank(int(co[0]))[: o.N]
        Fg = g.read(int(cg[0]))[: o.N]
        say("   max|dF|", float(np.abs(Fo - Fg).max()))
        assert np.abs(Fo - Fg).max() <= 1e-4
    g.close()
    o.close()


# non-default numerics and fixture options (water/coupled_gpu.gd:9-21 flags;
# height normals height_normal.glslinc:6-34, interior fast path
# coupled_motion.glslinc:119, linear air extension coupled_motion.glslinc:214-229)
OPTION_SETS = {
    "height_interior_linear_air": dict(height_normals=True, transport_interior=True, linear_air=True),
    "face_cfl": dict(max_face_cfl=True),
    "uncached_no_band": dict(transport_cache=False, active_advection=False),
}


@pytest.mark.parametrize("name", sorted(OPTION_SETS))
def test_batch_groups_nondefault_options(gpu, name):
    pair = P.Pair(narrow_flume(**SMALL, **OPTION_SETS[name]))
    run_batch(pair, where=name)
    pair.close()


# initial shapes 1-6 of op 40 (coupled_motion.glslinc:234-270): 1 dam-break
# block, 2 block over a floor layer, 3 uniform 0.25 m/s layer, 4 two colliding
# blocks, 5 linear wave (pressure_core.glslinc:54-60 surface), 6 flat tank with
# right absorption (coastal_target.glslinc:30, coastal_pass.glslinc:72)
FIXTURES = {1: dict(), 2: dict(), 3: dict(), 4: dict(), 5: dict(), 6: dict(slope=0.0, right_absorption=1.0)}


@pytest.mark.parametrize("shape", sorted(FIXTURES))
def test_fixture_shapes(gpu, shape):
    pair = P.Pair(narrow_flume(**SMALL, shape=shape, **FIXTURES[shape]))
    rep = P.compare_named(pair.o, pair.g)
    # shape 5's initial velocities come from the linear stream function
    # (sinh, cos, pow: coupled_motion.glslinc:203-207), the other shapes' are rational
    exact = ("img_kinds",) if shape == 5 else ("bankA", "bankB", "geo", "work", "motion", "img_kinds")
    for name in exact:
        assert rep[name][0] == 0, (shape, name, rep[name])
    kinds_o, kinds_g = pair.o.get(P.O.FIELD_GEO)[1::2], pair.g.read(P.S.FIELD_GEO)[1::2]
    assert np.array_equal(kinds_o, kinds_g), shape          # cell flags bit-exact in every shape
    for name in ("bankA", "bankB"):
        assert rep[name][2] < P.TRANSCENDENTAL_REL, (shape, name, rep[name])
    run_batch(pair, where=f"shape{shape}", transcendental_all=shape == 5)
    pair.close()


@pytest.mark.xfail(reason="periodic x with the coastal inlet: the level-0 operator is not symmetric there and the "
                          "reference's authoritative-triangle Galerkin scatter and this build's gather-and-mirror "
                          "Galerkin differ by about 4 % on coarse rows (the reference runs its batch program with "
                          "periodic_x false, water/lab_runner.gd:26)", strict=False)
def test_batch_groups_periodic_x(gpu):
    pair = P.Pair(narrow_flume(**SMALL, periodic_x=True))
    run_batch(pair, where="periodic_x")
    pair.close()


# ---- forced solver states: the rejection, retry and defect paths -------------
# (coupled_motion.glslinc:351-355; pressure_pass.glslinc:20-24, 78-103, 138)

def _set(**kv):
    def f(c):
        for k, v in kv.items():
            c[int(k[1:])] = v
    return f


@pytest.mark.parametrize("code,retries", [(20, 0), (20, 1), (20, 2), (10, 0), (11, 1), (12, 2), (4, 0), (21, 0), (22, 0)])
def test_retry_op43(gpu, code, retries):
    """op 43: a recoverable failure (CFL 20, VOF bounds 10-12) with fewer than
    2 retries rolls back: the trial closes, the dt cap halves; anything else,
    or an exhausted budget, stays failed and the trial start (op 1) opens
    nothing."""
    pair = P.Pair(narrow_flume(**SMALL))
    pair.o.set_quota(5.0)
    pair.g.request(5)
    dt = pair.o.control[20]
    rep = pair.group((43,), _set(c7=code, c29=retries, c4=1, c5=1, c20=dt * 0.75))
    P.check_group((43,), rep)
    c = pair.o.control
    retried = (code == 20 or 10 <= code <= 12) and retries < 2
    assert (c[7] == 0) == retried and (c[29] == retries + 1) == retried
    if retried:
        assert c[30] == np.float32(dt * 0.75) * np.float32(0.5) and c[33] == code
    rep = pair.group((1,))
    P.check_group((1,), rep)
    assert (pair.o.control[5] == 1) == retried    # a new trial opened
    pair.close()


def _advance_to(pair, stop, quota=1.0):
    """Run groups on identical inputs up to (not including) group index `stop`."""
    groups = P.batch_groups()
    pair.o.set_quota(quota)
    pair.g.request(int(quota))
    for k in range(stop):
        rep = pair.group(groups[k])
        P.check_group(groups[k], rep)
    return groups


ACCEPT_CASES = {
    "natural": None,
    "defect_request": _set(c61=0, c58=0),
    "defect_last": _set(c61=0, c58=3),
    "defects_exhausted": _set(c61=0, c58=4),
    "injected": _set(c19=1, c61=1),
    "mass_defect": None,    # c[26] moved off the ledger closure below
    "cfl": _set(c61=1, c27=0.46),
}


@pytest.mark.parametrize("case", sorted(ACCEPT_CASES))
def test_acceptance_branches(gpu, case):
    """ops 12/13 after one projected trial, each branch of the acceptance
    state machine forced through c[]: a defect-correction request (c[57] = 1,
    at most 4), reject 4 (no true convergence after 4 defects, or the op 15
    injection), mass defect 21, CFL 20, acceptance with the bank flip."""
    pair = P.Pair(narrow_flume(**SMALL))
    groups = _advance_to(pair, P.batch_groups().index((12, 13)))
    mut = ACCEPT_CASES[case]
    if case == "mass_defect":
        def mut(c):
            c[61] = 1
            c[26] = c[26] * 1.01
    if case == "injected":
        pair.group((15,))
    rep = pair.group((12, 13), mut)
    P.check_group((12, 13), rep, case)
    c = pair.o.control
    say(case, "c[2,3,6,7,57,58]", c[[2, 3, 6, 7, 57, 58]])
    expect = {"defect_request": (7, 0, 57, 1), "defect_last": (7, 0, 58, 4), "defects_exhausted": (7, 4),
              "injected": (7, 4), "mass_defect": (7, 21), "cfl": (7, 20)}.get(case)
    if case in ("mass_defect", "cfl") and not (c[14] < 1e-4 and c[15] < 1e20):
        expect = None    # the divergence check already asks for a defect correction
    if expect:
        for i in range(0, len(expect), 2):
            assert c[expect[i]] == expect[i + 1], (case, expect, c[:60])
    for grp in groups[groups.index((12, 13)) + 1:]:
        P.check_group(grp, pair.group(grp), case)
    pair.close()


def test_defect_correction_batch(gpu):
    """A defect-correction request carried into the next batch: op 63 opens the
    correction solve (c[57] 1 -> 2), op 64 saves and zeroes p, the hierarchy
    build and PCG set-up run with c[57] == 2, op 65 restores p, ops 62/66
    recompute the true residual; every group on identical inputs."""
    pair = P.Pair(narrow_flume(**SMALL))
    groups = _advance_to(pair, P.batch_groups().index((12, 13)))
    rep = pair.group((12, 13), _set(c61=0, c58=0))
    assert pair.o.control[57] == 1
    for grp in groups[groups.index((12, 13)) + 1:]:
        P.check_group(grp, pair.group(grp))
    ctl = run_batch(pair, where="defect")
    say("after the correction batch: c[2,6,7,57,58]", ctl[[2, 6, 7, 57, 58]])
    assert ctl[58] >= 1
    pair.close()


FAILURE_CASES = {
    # VOF sweep bounds (op 26: F outside [-1e-6, 1 + 1e-6] after the update -> 10 + sweep)
    "vof_bounds": ((25, 26), "F"),
    # forcing bounds (op 88: relaxed F outside [0, 1] -> 23)
    "forcing_bounds": ((87, 88), "Fc"),
    # compression coverage (op 46: a full cell outside the accepted pressure mask -> 22)
    "coverage": ((45, 46), "mask"),
}


@pytest.mark.parametrize("case", sorted(FAILURE_CASES))
def test_failure_codes(gpu, case):
    """Failure codes raised by the step's own checks, on states written into
    the fields: every kernel exits normally, c[7] and c[3] are compared."""
    pair = P.Pair(narrow_flume(**SMALL))
    grp, what = FAILURE_CASES[case]
    groups = P.batch_groups()
    _advance_to(pair, groups.index(grp))
    o = pair.o
    N = o.N
    wet = np.nonzero(o.get(P.O.FIELD_GEO)[1::2] > 0.5)[0]
    i = int(wet[len(wet) // 2])                     # a wet cell in the interior of the fluid
    c = o.control
    if what == "F":
        b = o.get(0 if int(c[0]) == 0 else 1)
        b[i] = 1.25
        o.set(0 if int(c[0]) == 0 else 1, b)
        pair.dirty.add(0 if int(c[0]) == 0 else 1)
    elif what == "Fc":
        cb = 1 - int(c[0])
        b = o.get(cb)
        b[i] = 1.5
        o.set(cb, b)
        pair.dirty.add(cb)
    else:
        m = o.get(P.O.FIELD_MOTION)
        acc = int(c[0])
        # pressure_mask_base(acc) = 4 N + 6 faces + acc N - the layout of coupled_motion.glslinc:8-14
        base = 2 * N + 6 * o.faces + acc * N
        full = np.nonzero(o.bank(acc)[:N] > 0.5)[0]
        m[base + full[len(full) // 2]] = 0
        o.set(P.O.FIELD_MOTION, m)
        pair.dirty.add(P.O.FIELD_MOTION)
    rep = pair.group(grp)
    P.check_group(grp, rep, case)
    c = o.control
    say(case, "c[3], c[7]", c[3], c[7])
    assert c[7] == {"vof_bounds": 10 + int(c[21]) - 1, "forcing_bounds": 23, "coverage": 22}[case]
    pair.close()


def test_mg_invalid_coefficients(gpu):
    """Hierarchy audit: a negative or non-finite fine coefficient stops the
    trial with code 34 (pressure_hierarchy_state.glsl:19-67)."""
    pair = P.Pair(narrow_flume(**SMALL))
    groups = P.batch_groups()
    _advance_to(pair, groups.index(("build",)))
    w = pair.o.get(P.O.FIELD_WORK)
    N = pair.o.N
    wet = np.nonzero(pair.o.get(P.O.FIELD_GEO)[1::2] > 0.5)[0]
    w[6 * N + wet[len(wet) // 2]] = -1.0
    pair.o.set(P.O.FIELD_WORK, w)
    pair.dirty.add(P.O.FIELD_WORK)
    rep = pair.group(("build",))
    P.check_group(("build",), rep)
    assert pair.o.control[7] == 34
    pair.close()


@pytest.mark.parametrize("fail_sweep", [False, True])
def test_batch_program_matches_ops(gpu, tmp_path, fail_sweep):
    """The batch program (one hipGraph replay: fused finalisers, the VOF
    sweeps' in-order ledger sums added by op 12's finaliser) against the same
    batches launched op by op through wv_op, on one device and identical
    inputs: every field bit-identical after every batch, up to the accepted
    step.  F = 1 + 5e-7 in one wet cell gives the sweeps nonzero clamp
    ledgers; fail_sweep: F = 1.25 in another makes sweep 0 fail (code 10), so
    the ledger sums of a failed trial are added too
    (coupled_motion.glslinc:402-408)."""
    from waves_amd import solver as S
    s = narrow_flume(**SMALL)
    a = S.Solver(s)
    a.start()
    for _ in range(2):
        a.step()
    path = str(tmp_path / "state.wvs")
    a.save_state(path)
    b = S.Solver(s)
    b.load_state(path)
    c = a.control
    acc = int(c[0])
    F = a.read(acc)
    wet = np.nonzero(a.read(S.FIELD_GEO)[1::2] > 0.5)[0]
    F[int(wet[len(wet) // 3])] = np.float32(1 + 5e-7)
    if fail_sweep:
        F[int(wet[2 * len(wet) // 3])] = 1.25
    a.write(acc, F)
    b.write(acc, F)
    target = int(c[2]) + 1
    a.request(target)
    b.request(target)
    for batch in range(4):
        a.frame(1)
        ca = a.poll(wait=True)
        for grp in P.batch_groups(s.iterations_per_batch):
            P.run_gpu(b, grp)
        cb = b.control
        say("fail_sweep", fail_sweep, "batch", batch, "c[7]", ca[7], cb[7], "c[22..25]", ca[22:26], cb[22:26],
            "accepted", ca[2], cb[2])
        for f in P.STATE_FIELDS:
            x, y = a.read(f), b.read(f)
            nd = int((x.view(np.uint32) != y.view(np.uint32)).sum())
            assert nd == 0, (batch, P.NAMES[f], nd)
        if ca[2] >= target or ca[7] > 0:
            break
    assert ca[7] == (10 if fail_sweep else 0)
    assert fail_sweep or ca[2] == target
    assert ca[22] != 0 or ca[24] != 0   # the clamp ledgers were formed (c[22] moves into c[24] on acceptance)
    a.close()
    b.close()
