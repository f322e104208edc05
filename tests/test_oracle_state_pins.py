"""Pins of the oracle's VOF transport, CFL rate, ghost-fluid coefficients,
relaxation, true residual, hierarchy audit and trial state machine, each by a
property the mathematics or the reference's documented rules fix (the
reference ships no tests or golden vectors), chosen so that a plausible
restatement mistake fails it:

  * VOF (coupled_motion.glslinc:375-408): a flat interface carried by a
    uniform vertical velocity moves by exactly C = v dt / h cells over the
    three split sweeps (a flipped flux sign, a wrong donor slab or a dropped
    divergence correction moves it elsewhere); a flooded box under a random
    velocity field stays exactly full (the divergence correction cancels the
    net fluxes); a mixed donor's flux equals |C| times the FP64
    inclusion-exclusion volume of its plane clipped to the swept slab.
  * CFL (coupled_motion.glslinc:189-202, 358-369): dt = min(dt_max, 0.45 /
    rate) with rate = |u|/hx + (|v| + g dt_max)/hy + |w|/hz for a uniform
    field, the largest term for the face-CFL option, the retry cap.
  * coefficients (pressure_core.glslinc:79-97): the ghost-fluid operator
    annihilates a hydrostatic (linear) pressure that vanishes on a flat
    surface cutting the cells at any height, and on the open top half a cell
    above the last row; solid and ocean-side faces carry no coefficient.
  * relaxation (coastal_pass.glslinc:13-59): F + (1 - exp(-r dt))(target - F)
    in FP64 with the quintic ramp of the rate and the still-water target.
  * true residual (coupled_motion.glslinc:271-284): ||1000/dt div u||^2 of one
    known face velocity.
  * hierarchy audit (pressure_hierarchy_state.glsl:19-67,
    water/pressure_hierarchy_state.gd:33-41): unchanged operator -> no
    rebuild; changed coefficient bits -> numerical rebuild only after `lag`
    trials in the lagged mode, at once in the structure mode; a flipped cell
    kind or face kind -> structural rebuild; an invalid coefficient -> 34.
  * trial state machine (coupled_motion.glslinc:351-355,
    pressure_pass.glslinc:20-24, 78-103, 138): retries halve the dt cap at
    most twice; acceptance branches by c[].
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O
from waves_amd.settings import Settings

RNG = np.random.default_rng(7)
SPACING = (0.025, 0.025, 0.05)


def box(n=(16, 24, 4), **kw):
    base = dict(n=n, spacing=SPACING, origin=(0.5, -0.3, -0.2), slope=0.0, amplitude=0.0,
                perturbation=0.0, history_capacity=64)
    base.update(kw)
    return Settings(**base)


def centres(o):
    nx, ny, nz = (int(v) for v in o.settings.n)
    h = np.asarray(o.settings.spacing, np.float64)
    org = np.asarray(o.settings.origin, np.float64)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    q = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    return q, org + (q + 0.5) * h


def face_slices(o):
    N = o.N
    return (slice(N, N + o.NU), slice(N + o.NU, N + o.NU + o.NV), slice(N + o.NU + o.NV, N + o.faces))


def set_control(o, **kv):
    c = o.control
    for k, v in kv.items():
        c[int(k[1:])] = v
    o.set(O.FIELD_CONTROL, c)
    return c


def write_bank(o, F=None, U=None, V=None, W=None, bank=0):
    b = o.bank(bank)
    if F is not None:
        b[: o.N] = F
    for comp, sl in zip((U, V, W), face_slices(o)):
        if comp is not None:
            b[sl] = comp
    o.set(O.FIELD_BANK_A if bank == 0 else O.FIELD_BANK_B, b)


def vof_step(o, dt):
    """The three split sweeps of one trial (ops 22, then 23/24/25/26 per sweep)."""
    set_control(o, c2=0, c5=1, c7=0, c20=dt, c21=0, c0=0, c4=1)
    o.op(22)
    for _ in range(3):
        for op in (23, 24, 55, 56, 25, 26):
            o.op(op)
    return o.bank(1)[: o.N]


# ---- VOF -------------------------------------------------------------------

@pytest.mark.parametrize("courant,level", [(0.25, 0.3), (0.5, 0.7), (0.75, 0.4), (-0.25, 0.6)])
def test_flat_interface_translates_by_courant_fraction(courant, level):
    s = box()
    o = O.Oracle(s)
    o.initial()
    q, xc = centres(o)
    h = SPACING[1]
    y0 = 10 * h + level * h                      # interface height above the bottom, in the box frame
    ybot = q[:, 1] * h
    F = np.clip((y0 - ybot) / h, 0, 1).astype(np.float32)
    dt = 1.0 / 120
    v = courant * h / dt
    nx, ny, nz = (int(a) for a in s.n)
    V = np.full((nz, ny + 1, nx), v, np.float32)
    V[:, 0, :] = 0                                # the bed face
    V[:, ny, :] = 0                               # the top face
    write_bank(o, F=F, U=np.zeros(o.NU, np.float32), V=V.ravel(), W=np.zeros(o.NW, np.float32))
    Fn = vof_step(o, dt).reshape(nz, ny, nx).astype(np.float64)
    expect = np.clip((y0 + courant * h - ybot) / h, 0, 1).reshape(nz, ny, nx)
    err = np.abs(Fn - expect)[:, :, 1:]          # column 0 samples the ocean ghost in its gradient
    assert err.max() <= 2e-6, err.max()
    assert o.control[7] == 0


def test_flooded_box_stays_full():
    s = box(n=(12, 10, 4))
    o = O.Oracle(s)
    o.initial()
    nx, ny, nz = (int(a) for a in s.n)
    dt = 1.0 / 120
    lim = 0.3 * SPACING[1] / dt
    U = RNG.uniform(-lim, lim, (nz, ny, nx + 1)).astype(np.float32)
    U[:, :, 0] = 0
    U[:, :, nx] = 0
    V = RNG.uniform(-lim, lim, (nz, ny + 1, nx)).astype(np.float32)
    V[:, 0, :] = 0
    V[:, ny, :] = 0
    W = RNG.uniform(-lim, lim, (nz + 1, ny, nx)).astype(np.float32)
    W[nz] = W[0]
    write_bank(o, F=np.ones(o.N, np.float32), U=U.ravel(), V=V.ravel(), W=W.ravel())
    Fn = vof_step(o, dt)
    c = o.control
    assert c[7] == 0
    assert np.all(Fn == 1.0), np.abs(Fn - 1).max()
    # the clamps that kept it exact are booked in the roundoff ledgers
    assert c[23] >= abs(c[22])


def test_mixed_donor_flux_is_the_clipped_plane_volume():
    s = box(n=(12, 16, 4))
    o = O.Oracle(s)
    o.initial()
    q, xc = centres(o)
    h = np.asarray(SPACING)
    # a tilted plane, liquid where n.(x - p0) <= 0: F from its FP64 volume in each cell
    nrm = np.array([0.3, 1.0, 0.15])
    base = np.array([0, 8.37, 0]) * h
    F = np.array([exact_volume(nrm * h, float(np.dot(nrm, base - qq * h))) for qq in q], np.float32)
    dt = np.float32(1.0 / 120)
    nx, ny, nz = (int(a) for a in s.n)
    v = np.float32(0.3 * h[1] / dt)
    V = np.zeros((nz, ny + 1, nx), np.float32)
    V[:, 1:ny, :] = v
    write_bank(o, F=F, U=np.zeros(o.NU, np.float32), V=V.ravel(), W=np.zeros(o.NW, np.float32))
    set_control(o, c2=1, c5=1, c7=0, c20=dt, c21=0, c0=0, c4=1)   # sweep axis (c[2] + c[21]) % 3 = y
    o.op(22)
    o.op(23)
    o.op(24)
    courant = float(v * dt / np.float32(h[1]))      # face_courant in FP32 (coupled_motion.glslinc:47-48)
    planes = o.get(O.FIELD_PLANE).reshape(-1, 4).astype(np.float64)
    flux = o.get(O.FIELD_MOTION)[2 * o.N + o.NU: 2 * o.N + o.NU + o.NV].reshape(nz, ny + 1, nx)
    checked = 0
    for z in range(nz):
        for y in range(1, ny):
            for x in range(1, nx - 1):
                d = x + nx * ((y - 1) + ny * z)          # donor below the face (v > 0)
                f = float(F[d])
                if not (0 < f < 1):
                    continue
                nq = planes[d, :3]
                alpha = planes[d, 3]
                ext = np.array([1.0, courant, 1.0])
                low = np.array([0.0, 1 - courant, 0.0])
                want = courant * exact_volume(nq * ext, alpha - float(np.dot(nq, low)))
                assert abs(float(flux[z, y, x]) - want) <= 2e-6, (z, y, x, flux[z, y, x], want)
                checked += 1
    assert checked >= 20


def exact_volume(n, alpha):
    """FP64 volume of {x in [0,1]^3 : n.x <= alpha} by inclusion-exclusion over
    the vertices of the axes with a nonzero component (k of them):
    sum_v (-1)^|v| max(alpha' - m.v, 0)^k / (k! prod m)."""
    n = np.asarray(n, np.float64)
    a = alpha - np.minimum(n, 0).sum()           # reflect negative components: x_i -> 1 - x_i
    m = np.abs(n)[np.abs(n) > 0]
    k = len(m)
    if k == 0:
        return 1.0 if a >= 0 else 0.0
    total = 0.0
    for v in itertools.product((0, 1), repeat=k):
        t = a - float(np.dot(m, v))
        if t > 0:
            total += (-1) ** sum(v) * t ** k
    return min(max(total / (math.factorial(k) * float(np.prod(m))), 0.0), 1.0)


# ---- CFL -------------------------------------------------------------------

@pytest.mark.parametrize("face_cfl", [False, True])
def test_cfl_rate_of_uniform_velocity(face_cfl):
    s = box(max_face_cfl=face_cfl)
    o = O.Oracle(s)
    o.initial()
    u, v, w = 1.3, -2.1, 0.9                      # fast enough that 0.45 / rate < dt_max
    write_bank(o, U=np.full(o.NU, u, np.float32), V=np.full(o.NV, v, np.float32), W=np.full(o.NW, w, np.float32))
    set_control(o, c5=1, c7=0, c29=0, c0=0)
    o.op(20)
    o.op(21)
    dtmax = float(np.float32(s.dt))
    hx, hy, hz = SPACING
    terms = (abs(u) / hx, (abs(v) + 9.81 * dtmax) / hy, abs(w) / hz)
    rate = max(terms) if face_cfl else sum(terms)
    want = min(dtmax, 0.45 / rate)
    assert want < dtmax
    assert abs(o.control[20] - want) <= 1e-6 * want, (o.control[20], want)
    # a retry caps dt at half the rejected one (c[30], coupled_motion.glslinc:351-355, 364)
    set_control(o, c29=1, c30=want / 4)
    o.op(20)
    o.op(21)
    assert o.control[20] == np.float32(want / 4)


# ---- ghost-fluid coefficients ----------------------------------------------

@pytest.mark.parametrize("theta", [0.5, 0.37, 0.9, 0.03])
def test_coefficients_annihilate_hydrostatic_pressure(theta):
    """A flat surface between rows j and j+1 at theta of the way up from the
    wet centre: the signed distance phi = y - eta in every cell, kinds from
    its sign; then A p = 0 for p = rho g (eta - y) on every wet cell not
    touching the bed, and on full columns with the zero level half a cell
    above the top row."""
    s = box(n=(10, 12, 4))
    o = O.Oracle(s)
    o.initial()
    q, xc = centres(o)
    h = SPACING[1]
    j = 7
    eta = xc[:, 1].reshape(-1)[0] + (j + theta) * h      # centre of row j plus theta h
    nx, ny, nz = (int(a) for a in s.n)
    phi = (xc[:, 1] - eta).astype(np.float32)
    geo = np.stack([phi, np.where(phi < 0, 1.0, 0.0).astype(np.float32)], 1).ravel()
    o.set(O.FIELD_GEO, geo)
    set_control(o, c4=0, c5=0)
    o.op(16)
    work = o.get(O.FIELD_WORK).reshape(14, o.N).astype(np.float64)
    p = np.where(phi < 0, 9810.0 * (eta - xc[:, 1]), 0.0)
    Ap = apply_operator(work, p, phi < 0, s.n)
    wet = (phi < 0) & (q[:, 1] >= 1) & (q[:, 0] >= 1) & (q[:, 0] <= nx - 2)
    scale = 9810.0 * h / h ** 2
    assert np.abs(Ap[wet]).max() <= 1e-4 * scale, np.abs(Ap[wet]).max() / scale
    # faces towards the bed (row 0 has the box bottom below) and the ocean side carry nothing
    k = work[6:12]
    bottom = q[:, 1] == 0
    assert np.all(k[2][bottom] == 0)          # -y face of the first row: y < 0 is solid
    assert np.all(k[0][q[:, 0] == 0] == 0)    # -x face of column 0: the ocean side (x < 0)
    assert np.all(k[1][q[:, 0] == nx - 1] == 0)   # +x face of the last column: solid (x >= nx)
    # a flooded column: the open top is a Dirichlet face at half a cell
    phi_full = (xc[:, 1] - (xc[:, 1].max() + 0.5 * h)).astype(np.float32) - 1.0
    geo = np.stack([phi_full, np.ones(o.N, np.float32)], 1).ravel()
    o.set(O.FIELD_GEO, geo)
    o.op(16)
    work = o.get(O.FIELD_WORK).reshape(14, o.N).astype(np.float64)
    top = xc[:, 1].max() + 0.5 * h
    p = 9810.0 * (top - xc[:, 1])
    Ap = apply_operator(work, p, np.ones(o.N, bool), s.n)
    inner = (q[:, 1] >= 1) & (q[:, 0] >= 1) & (q[:, 0] <= nx - 2)
    assert np.abs(Ap[inner]).max() <= 1e-4 * scale


def apply_operator(work, p, wet, n):
    nx, ny, nz = (int(a) for a in n)
    P = np.where(wet, p, 0.0).reshape(nz, ny, nx)
    out = np.zeros_like(P)
    face = 0
    for ax in (2, 1, 0):
        for sign in (-1, 1):
            if ax == 0:
                nb = np.roll(P, -sign, axis=0)
            else:
                nb = np.zeros_like(P)
                src = [slice(None)] * 3
                dst = [slice(None)] * 3
                if sign < 0:
                    dst[ax], src[ax] = slice(1, None), slice(None, -1)
                else:
                    dst[ax], src[ax] = slice(None, -1), slice(1, None)
                nb[tuple(dst)] = P[tuple(src)]
            out += work[6 + face].reshape(nz, ny, nx) * (P - nb)
            face += 1
    return np.where(wet.reshape(nz, ny, nx), out, 0.0).ravel()


# ---- relaxation ------------------------------------------------------------

def smoothstep5(x):
    x = np.clip(x, 0, 1)
    return x ** 3 * (10 - 15 * x + 6 * x ** 2)


def test_relaxation_of_fraction_and_faces():
    s = box(n=(40, 16, 2), zone_length=0.5)
    o = O.Oracle(s)
    o.initial()
    q, xc = centres(o)
    dt = 1.0 / 150
    F = RNG.uniform(0, 1, o.N).astype(np.float32)
    U = RNG.normal(0, 0.2, o.NU).astype(np.float32)
    write_bank(o, F=F, U=U, bank=1)
    set_control(o, c0=0, c4=1, c5=1, c7=0, c20=dt, c1=0.0)
    o.op(80)
    Fn = o.bank(1)[: o.N].astype(np.float64)
    x0, zone, rate0 = s.origin[0], s.zone_length, s.damping_rate
    r = rate0 * (1 - smoothstep5((xc[:, 0] - x0) / zone))
    target = np.clip(0.5 - xc[:, 1] / SPACING[1], 0, 1)      # the still surface at y = 0 (amplitude 0)
    want = F + (1 - np.exp(-r * dt)) * (target - F)
    assert np.abs(Fn - want).max() <= 2e-6
    # faces: u relaxed towards the (zero) wave velocity, the inflow face takes the target
    o.op(82)
    Un = o.bank(1)[o.N: o.N + o.NU].astype(np.float64)
    nx, ny, nz = (int(a) for a in s.n)
    qx = np.arange(o.NU) % (nx + 1)
    xf = x0 + qx * SPACING[0]
    rf = rate0 * (1 - smoothstep5((xf - x0) / zone))
    wantU = U * np.exp(-rf * dt)                  # the wave velocity is zero at amplitude 0
    wantU[qx == 0] = 0.0
    wall = qx == nx
    assert np.abs(Un - wantU)[~wall].max() <= 2e-6


# ---- true residual ----------------------------------------------------------

def test_true_residual_of_one_face_velocity():
    s = box(n=(8, 8, 2))
    o = O.Oracle(s)
    o.initial()
    nx, ny, nz = (int(a) for a in s.n)
    geo = o.get(O.FIELD_GEO).reshape(-1, 2)
    geo[:, 1] = 1.0                               # every cell wet
    o.set(O.FIELD_GEO, geo.ravel())
    dt, u = 1.0 / 200, 0.04
    U = np.zeros((nz, ny, nx + 1), np.float32)
    U[0, 3, 4] = u                                # one interior x face
    write_bank(o, U=U.ravel(), V=np.zeros(o.NV, np.float32), W=np.zeros(o.NW, np.float32), bank=1)
    b2 = 7.5
    set_control(o, c0=0, c4=1, c6=1, c7=0, c20=dt, c59=b2)
    o.op(62)
    o.op(66)
    c = o.control
    r = 1000 / dt * u / SPACING[0]                # the two cells beside the face: +-r
    r2 = 2 * r * r
    assert abs(c[60] - math.sqrt(r2 / b2)) <= 1e-6 * math.sqrt(r2 / b2)


# ---- hierarchy audit --------------------------------------------------------

def audit_state(o):
    return o.get(O.FIELD_AUDIT, np.uint32)


def built_oracle(mode=2, lag=8):
    s = box(n=(16, 24, 4), mg_maintenance=mode, mg_lag=lag)
    o = O.Oracle(s)
    o.initial()
    set_control(o, c4=1, c5=1, c7=0)
    o.mg_record(True)                             # the first build: no hierarchy yet -> full rebuild
    st = audit_state(o)
    assert st[8] == 1 and st[9] == 1 and st[0] == 1
    return o


def rebuild_flags(o):
    set_control(o, c4=1, c5=1, c7=0)
    o.mg_record(True)
    st = audit_state(o)
    return int(st[8]), int(st[9]), int(st[10]), int(st[3])


def test_audit_unchanged_operator_keeps_the_hierarchy():
    o = built_oracle()
    for _ in range(3):
        assert rebuild_flags(o) == (0, 0, 0, 0)


def test_audit_lagged_numerical_rebuild():
    lag = 4
    o = built_oracle(mode=2, lag=lag)
    work = o.get(O.FIELD_WORK).reshape(14, -1)
    geo = o.get(O.FIELD_GEO).reshape(-1, 2)
    i = int(np.flatnonzero((geo[:, 1] > 0.5) & (work[6] > 0))[5])
    work[6, i] = np.nextafter(work[6, i], np.float32(np.inf))  # one coefficient one ulp off
    o.set(O.FIELD_WORK, work.ravel())
    seen = [rebuild_flags(o) for _ in range(lag)]
    # ages 1 .. lag-1 without a rebuild, then the numerical rebuild at age lag
    assert [f[0] for f in seen] == [0] * (lag - 1) + [1], seen
    assert [f[2] for f in seen] == list(range(1, lag)) + [0], seen
    assert all(f[1] == 0 for f in seen) and all(f[3] == 1 for f in seen)


def test_audit_structure_mode_rebuilds_numerically_at_once():
    o = built_oracle(mode=1)
    work = o.get(O.FIELD_WORK).reshape(14, -1)
    geo = o.get(O.FIELD_GEO).reshape(-1, 2)
    i = int(np.flatnonzero((geo[:, 1] > 0.5) & (work[6] > 0))[3])
    work[6, i] *= np.float32(1.5)
    o.set(O.FIELD_WORK, work.ravel())
    assert rebuild_flags(o)[:2] == (1, 0)


@pytest.mark.parametrize("what", ["kind", "face"])
def test_audit_structural_rebuild(what):
    o = built_oracle()
    work = o.get(O.FIELD_WORK).reshape(14, -1)
    geo = o.get(O.FIELD_GEO).reshape(-1, 2)
    wet = np.flatnonzero(geo[:, 1] > 0.5)
    if what == "kind":
        geo[wet[-1], 1] = 0.0                      # a wet cell turns dry
        o.set(O.FIELD_GEO, geo.ravel())
    else:
        i = int(wet[np.flatnonzero(work[7][wet] > 0)[0]])
        work[7, i] = 0.0                           # a face kind changes (open -> none)
        o.set(O.FIELD_WORK, work.ravel())
    f = rebuild_flags(o)
    assert f[0] == 1 and f[1] == 1, f


def test_audit_invalid_coefficient_is_failure_34():
    o = built_oracle()
    work = o.get(O.FIELD_WORK).reshape(14, -1)
    work[8, 10] = np.float32(np.nan)
    o.set(O.FIELD_WORK, work.ravel())
    set_control(o, c4=1, c5=1, c7=0)
    o.mg_record(True)
    assert o.control[7] == 34


# ---- trial state machine ----------------------------------------------------

@pytest.mark.parametrize("code", [20, 10, 11, 12, 21, 4, 23])
def test_retry_budget(code):
    o = O.Oracle(box(n=(8, 8, 2)))
    o.initial()
    recoverable = code == 20 or 10 <= code <= 12
    dt = 0.004
    for attempt in range(4):
        set_control(o, c7=code, c20=dt, c4=1, c5=1)
        o.op(43)
        c = o.control
        if recoverable and attempt < 2:
            assert c[7] == 0 and c[29] == attempt + 1 and c[30] == np.float32(dt) * np.float32(0.5)
            assert c[33] == code and c[34] == np.float32(dt) and c[4] == 0 and c[5] == 0
        else:
            assert c[7] == code and c[4] == 1            # stays failed: no retry left or not recoverable


def test_trial_start_quota():
    o = O.Oracle(box(n=(8, 8, 2)))
    o.initial()
    set_control(o, c2=3, c4=0, c7=0, c32=5)
    o.set_quota(3)
    o.op(1)
    assert o.control[4] == 0 and o.control[5] == 0               # quota of accepted steps met
    o.set_quota(4)
    o.op(1)
    c = o.control
    assert c[4] == 1 and c[5] == 1 and c[32] == 6 and c[6] == 0 and c[8] == 0
    o.op(1)                                                       # a trial already open: no new one
    assert o.control[5] == 0 and o.control[32] == 6
    set_control(o, c4=0, c1=2.0)
    o.set_quota(-2.0)                                             # a deadline of 2 s, reached
    o.op(1)
    assert o.control[4] == 0


def partials_for_acceptance(o, div, speed, err, volume):
    """Write the diagnostics partials op 13 reduces (first group holds all)."""
    G = (o.N + 127) // 128
    part = np.zeros((G, 4), np.float32)
    part[0] = (div, speed, err, volume)
    o.set(O.FIELD_PARTIAL, part.ravel())


@pytest.mark.parametrize("case", ["accept", "defect", "defect_last", "defects_exhausted", "injected", "mass", "cfl", "nan"])
def test_acceptance_state_machine(case):
    o = O.Oracle(box(n=(8, 8, 2)))
    o.initial()
    V0 = 1.25
    c = set_control(o, c0=0, c1=1.0, c2=10, c3=0, c4=1, c6=1, c7=0, c19=0, c20=0.004, c26=V0, c27=0.3, c58=0,
                    c61=1, c22=0, c24=0, c39=0, c40=0)
    for k in (64, 65, 66, 72, 73, 74):
        c[k] = 0
    o.set(O.FIELD_CONTROL, c)
    div, speed, vol = 1e-5, 0.5, V0
    if case == "defect":
        set_control(o, c61=0)
    if case == "defect_last":
        set_control(o, c61=0, c58=3)
    if case == "defects_exhausted":
        set_control(o, c61=0, c58=4)
    if case == "injected":
        o.op(15)
    if case == "mass":
        vol = V0 * 1.01
    if case == "cfl":
        set_control(o, c27=0.46)
    if case == "nan":
        speed = 1e30
    partials_for_acceptance(o, div, speed, 0.0, vol)
    o.op(13)
    c = o.control
    expect = {"accept": dict(c0=1, c2=11, c4=0, c7=0), "defect": dict(c57=1, c58=1, c6=0, c7=0, c2=10),
              "defect_last": dict(c57=1, c58=4, c7=0),
              "defects_exhausted": dict(c7=4, c3=1, c2=10), "injected": dict(c7=4, c3=1),
              "mass": dict(c7=21, c3=1), "cfl": dict(c7=20, c3=1, c2=10), "nan": dict(c7=4, c3=1)}[case]
    for k, v in expect.items():
        assert c[int(k[1:])] == v, (case, k, c[int(k[1:])], v)
    if case == "accept":
        assert c[1] == np.float32(1.0) + np.float32(0.004) and c[90] == np.float32(0.004)
        assert c[17] == np.float32(V0) and abs(c[28]) < 1e-6


# ---- height-function normals, compression audit, forcing ledgers -------------

@pytest.mark.parametrize("sx,sz", [(0.3, 0.0), (0.25, 0.1), (-0.2, -0.15)])
def test_height_normals_are_exact_for_a_plane(sx, sz):
    """height_normal.glslinc:6-34: for a planar surface y = y0 + sx x + sz z the
    column sums of exact cell fractions are the plane's height at the column
    centre, so the central differences give the plane's normal exactly
    (-sx, 1, -sz); the Youngs gradient alone does not."""
    def normals(height_normals):
        s = box(n=(16, 24, 8), height_normals=height_normals)
        o = O.Oracle(s)
        o.initial()
        q, xc = centres(o)
        h = np.asarray(SPACING)
        y0 = 0.30
        nrm = np.array([-sx, 1.0, -sz])              # liquid where nrm.(x - p) <= 0, x local to the box
        F = np.array([exact_volume(nrm * h, float(y0 - np.dot(nrm, qq * h) + nrm[1] * 0)) for qq in q], np.float32)
        write_bank(o, F=F)
        set_control(o, c0=0, c4=0, c5=0, c7=0)
        o.op(17)
        planes = o.get(O.FIELD_PLANE).reshape(-1, 4).astype(np.float64)
        nx, ny, nz = (int(a) for a in s.n)
        inner = (F > 0) & (F < 1) & (q[:, 0] >= 1) & (q[:, 0] <= nx - 2) & (q[:, 2] >= 1) & (q[:, 2] <= nz - 2)
        want = nrm * h / np.abs(nrm * h).sum()
        return np.abs(planes[inner, :3] - want).max(), int(inner.sum())
    err_h, cnt = normals(True)
    err_y, _ = normals(False)
    assert cnt >= 20
    assert err_h <= 1e-5, err_h
    if sx != 0 and sz != 0:
        assert err_y > 1e-4, err_y          # the pin tells the two paths apart


@pytest.mark.parametrize("cut", [None, 7])
def test_compression_audit_obeys_gauss(cut):
    """ops 45/46 (coupled_motion.glslinc:341-350): sum over full cells (F > 1/2)
    of dt div u V telescopes to dt times the net flux out of the full region
    through its open faces (walls carry none); the F V sum is the liquid
    volume.  cut: columns x >= cut hold F = 0.45 (not full), so the full region
    is the sub-box x < cut.  Anisotropic spacing."""
    sp = (0.025, 0.02, 0.05)
    s = box(n=(12, 10, 4), spacing=sp)
    o = O.Oracle(s)
    o.initial()
    nx, ny, nz = (int(a) for a in s.n)
    hx, hy, hz = sp
    U = RNG.normal(0, 0.1, (nz, ny, nx + 1)).astype(np.float32)
    V = RNG.normal(0, 0.1, (nz, ny + 1, nx)).astype(np.float32)
    W = RNG.normal(0, 0.1, (nz + 1, ny, nx)).astype(np.float32)
    W[nz] = W[0]
    F = RNG.uniform(0.6, 1.0, (nz, ny, nx)).astype(np.float32)
    k = nx if cut is None else cut
    F[:, :, k:] = 0.45
    F = F.ravel()
    write_bank(o, F=F, U=U.ravel(), V=V.ravel(), W=W.ravel())
    motion = o.get(O.FIELD_MOTION)
    pm = 2 * o.N + 6 * o.faces                     # the accepted pressure mask (bank 0)
    motion[pm:pm + o.N] = 1
    o.set(O.FIELD_MOTION, motion)
    dt = 1.0 / 150
    set_control(o, c0=0, c5=1, c7=0, c20=dt)
    o.op(45)
    o.op(46)
    c = o.control
    U64, V64 = U.astype(np.float64), V.astype(np.float64)
    # open faces of the full region: the ocean side x = 0, its right side x = k
    # (a wall only at k = nx), the top y = ny (y = 0 is the bed)
    right = U64[:, :, k].sum() if k < nx else 0.0
    raw = dt * ((U64[:, :, k].sum() - U64[:, :, 0].sum()) * hy * hz + (V64[:, ny, :k].sum() - V64[:, 0, :k].sum()) * hx * hz)
    walled = dt * ((right - U64[:, :, 0].sum()) * hy * hz + V64[:, ny, :k].sum() * hx * hz)
    scale = dt * 0.1 * hx * hz * nx * nz * 4
    assert abs(c[38] - raw) <= 1e-4 * scale, (c[38], raw)
    assert abs(c[35] - walled) <= 1e-4 * scale, (c[35], walled)
    assert c[36] == 0 and c[7] == 0
    assert abs(c[37] - F.astype(np.float64).sum() * hx * hy * hz) <= 1e-6 * c[37]


def test_forcing_ledgers():
    """op 83/84: the momentum put in by the face relaxation is the mass times
    the relaxed velocity change; op 87/88: the relaxed F's extremes, failure 23
    outside [0, 1]."""
    s = box(n=(12, 10, 4))
    o = O.Oracle(s)
    o.initial()
    hx, hy, hz = SPACING
    F = RNG.uniform(0, 1, o.N).astype(np.float32)
    write_bank(o, F=F, bank=1)
    motion = o.get(O.FIELD_MOTION)
    fwd = 2 * o.N + o.faces                        # the forward field holds op 82's face changes
    du = np.float32(0.013)
    motion[fwd:fwd + o.NU] = du
    motion[fwd + o.NU:fwd + o.faces] = 0
    o.set(O.FIELD_MOTION, motion)
    set_control(o, c0=0, c5=1, c7=0, c20=1 / 150, c67=0, c68=0, c69=0)
    o.op(83)
    o.op(84)
    c = o.control
    want = float(du) * 1000 * F.astype(np.float64).sum() * hx * hy * hz
    assert abs(c[67] - want) <= 1e-5 * want and c[68] == 0 and c[69] == 0
    o.op(87)
    o.op(88)
    c = o.control
    # lanes past the last cell of the last 128-lane group reduce an F of 0
    # (coastal_pass.glslinc:51-58): they count in the minimum
    pad = o.N % 128 != 0
    assert c[88] == (min(F.min(), 0) if pad else F.min()) and c[89] == F.max() and c[7] == 0
    F[5] = 1.25
    write_bank(o, F=F, bank=1)
    o.op(87)
    o.op(88)
    assert o.control[7] == 23 and o.control[89] == np.float32(1.25)
