"""GPU parity: every op group of a batch on identical inputs, and short
trajectories, HIP solver vs CPU oracle (water/shaders/compute/*.glslinc semantics)."""
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
    rep = P.compare(o, g)
    for k, v in rep.items():
        say("initial", k, v)
    # rational ops are bit-exact; the probes evaluate sinh/cosh/tanh/cos/sin
    for name in ("bankA", "bankB", "work", "geo", "plane", "motion", "img_planes", "img_kinds"):
        assert rep[name][0] == 0, (name, rep[name])
    assert rep["probe"][2] < 1e-5
    c_o, c_g = o.control, g.control
    assert c_o[26] == c_g[26]  # initial volume, same reduction tree
    np.testing.assert_allclose(c_g[80:88], c_o[80:88], rtol=0, atol=0)


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
    rel = float(np.abs(ro - rg).max() / max(np.abs(ro).max(), 1e-30))
    say("  coarse rows rel diff", rel)
    assert rel < tol
    # the coarse sweeps address every row by delta + one periodic correction,
    # which needs the entries past a non-periodic edge to be exactly zero
    for r0, N, W, mask in P.coarse_edge_entries(o):
        assert np.all(rg[r0: r0 + W * N].reshape(W, N)[mask] == 0), ("edge entries", r0)


def _batch_report(s, quota=1.0, groups=None, check=True, mg_tol=1e-5):
    o, g = P.make_pair(s)
    P.start_pair(o, g)
    o.set_quota(quota)
    g.request(int(quota))
    worst = {}
    for grp in groups or P.batch_groups():
        P.sync_state(o, g)
        if any(x in ("cycle0", "cycle1") for x in grp):
            P.sync_mg(o, g)
        P.run_oracle(o, grp)
        P.run_gpu(g, grp)
        rep = P.compare(o, g, P.STATE_FIELDS)
        if "build" in grp:
            rep.update(P.compare(o, g, (P.O.FIELD_SNAPSHOT, P.O.FIELD_AUDIT)))
            mg_report(o, g, mg_tol)
        if any(x in ("cycle0", "cycle1") for x in grp):
            rep.update(P.compare(o, g, (P.O.FIELD_MG_VECTORS,)))
        bad = {k: v for k, v in rep.items() if v[0]}
        say("group", grp, "c[4..8]", o.control[4:9], bad if bad else "exact")
        for k, v in bad.items():
            key = (grp, k)
            worst[key] = v
    return o, g, worst


def test_batch_groups_on_identical_inputs(gpu):
    s = narrow_flume(**SMALL)
    o, g, worst = _batch_report(s)
    # every group whose arithmetic is rational must be bit-exact; groups that
    # evaluate transcendentals (coastal forcing, ocean ghost velocities) within 2e-5 relative
    for (grp, field), (nd, mx, rel) in worst.items():
        assert rel < 2e-5, (grp, field, nd, mx, rel)


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
    s = narrow_flume(**SMALL, **OPTION_SETS[name])
    o, g, worst = _batch_report(s)
    for (grp, field), (nd, mx, rel) in worst.items():
        assert rel < 2e-5, (name, grp, field, nd, mx, rel)

