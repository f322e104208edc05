"""Pins of the CPU oracle against what the mathematics fixes (the reference has
no tests, fixtures or golden vectors; its GLSL cannot run in this container).

Each check is chosen so that a plausible restatement mistake (a dropped term, a
wrong sign or index, a transposed operand) fails it:
  * cube volume vs an FP64 inclusion-exclusion closed form (plic_math.glslinc:3-23)
  * plane-constant bisection round trip and tie rules (plic_math.glslinc:24-37)
  * Stokes-2 probes vs the textbook FP64 formulas (coastal_target.glslinc, op 86)
  * the 7-point masked operator vs a dense brute-force matrix; SPD (pressure_core.glslinc:120-125)
  * Galerkin coarse operators vs a dense FP64 P^T A P (pressure_multigrid.glslinc:159-175)
  * Cholesky L^-1 vs numpy (pressure_mg_coarse.glslinc:35-49)
  * the V-cycle is a symmetric positive operator (pressure_multigrid.glslinc:191-234)
  * still water: the projected pressure is hydrostatic, velocities stay zero
  * a pressure solve agrees with an FP64 sparse direct solve of the same system
  * mass ledger closes over a few wave steps (pressure_pass.glslinc:91-95)
"""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from waves_amd.settings import Settings, dispersion_k, narrow_flume

RNG = np.random.default_rng(20261203)


def exact_cube_volume(n, alpha):
    """Volume of {x in [0,1]^3 : n.x <= alpha} in FP64 by inclusion-exclusion
    over the cube's vertices (all |n_i| > 0)."""
    n = np.asarray(n, np.float64)
    # reflect negative components: x_i -> 1 - x_i
    a = alpha - np.minimum(n, 0).sum()
    m = np.abs(n)
    total = 0.0
    for v in itertools.product((0, 1), repeat=3):
        t = a - float(np.dot(m, v))
        if t > 0:
            total += (-1) ** sum(v) * t ** 3
    return min(max(total / (6 * m[0] * m[1] * m[2]), 0.0), 1.0)


def test_cube_volume_matches_fp64_closed_form():
    worst = 0.0
    for _ in range(3000):
        n = RNG.normal(size=3)
        n[np.abs(n) < 0.05] = 0.05
        n = (n / np.abs(n).sum()).astype(np.float32)
        lo, hi = np.minimum(n, 0).sum(), np.maximum(n, 0).sum()
        alpha = np.float32(RNG.uniform(lo - 0.05, hi + 0.05))
        worst = max(worst, abs(O.cube_volume(n, alpha) - exact_cube_volume(n.astype(np.float64), float(alpha))))
    assert worst < 2e-6, worst


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_cube_volume_axis_aligned_and_edge_cases(axis):
    n = np.zeros(3, np.float32)
    n[axis] = 1.0
    for a in np.linspace(-0.2, 1.2, 29).astype(np.float32):
        assert O.cube_volume(n, a) == pytest.approx(min(max(float(a), 0.0), 1.0), abs=1e-6)
    n[axis] = -1.0  # n.x <= a  <=>  x_axis >= -a
    for a in np.linspace(-1.2, 0.2, 29).astype(np.float32):
        assert O.cube_volume(n, a) == pytest.approx(min(max(1.0 + float(a), 0.0), 1.0), abs=1e-6)
    # a two-component normal: a prism of the unit square (2-D closed form)
    n2 = np.zeros(3, np.float32)
    n2[axis] = 0.3
    n2[(axis + 1) % 3] = 0.7
    for a in np.linspace(0, 1, 11).astype(np.float32):
        x, y = 0.3, 0.7
        # area under x*u + y*v <= a in the unit square
        area = 0.0
        for u0, v0 in ((0, 0), (1, 0), (0, 1), (1, 1)):
            t = float(a) - x * u0 - y * v0
            if t > 0:
                area += (-1) ** (u0 + v0) * t * t
        area /= 2 * x * y
        assert O.cube_volume(n2, a) == pytest.approx(min(max(area, 0), 1), abs=2e-6)


def test_cube_volume_symmetries_and_monotonicity():
    for _ in range(500):
        n = RNG.normal(size=3).astype(np.float32)
        n /= np.abs(n).sum()
        centre = float(np.dot(n, [0.5, 0.5, 0.5]))
        assert O.cube_volume(n, np.float32(centre)) == pytest.approx(0.5, abs=2e-6)
        a = np.float32(RNG.uniform(-1, 1))
        assert O.cube_volume(n, a) + O.cube_volume(-n, -a) == pytest.approx(1.0, abs=3e-6)
        grid = np.linspace(-1.1, 1.1, 41).astype(np.float32)
        vals = [O.cube_volume(n, g) for g in grid]
        assert all(b >= a - 1e-7 for a, b in zip(vals, vals[1:]))


def test_inverse_volume_round_trip_and_tie_rules():
    for _ in range(800):
        n = RNG.normal(size=3).astype(np.float32)
        n /= np.abs(n).sum()
        f = np.float32(RNG.uniform(0.001, 0.999))
        alpha = O.inverse_volume(n, f)
        assert abs(O.cube_volume(n, np.float32(alpha)) - float(f)) < 1e-6
        h = np.float32(0.5)
        centre = (n[0] * h + n[1] * h) + n[2] * h  # float32, left to right
        if f < 0.5:
            assert alpha <= centre
        elif f > 0.5:
            assert alpha > centre  # geometry-derived wet flag agrees with F > 1/2
        assert O.inverse_volume(n, np.float32(0.5)) == pytest.approx(float(centre), abs=0)


def test_dispersion_relation():
    k = dispersion_k(1.4, 0.45)
    w = 2 * math.pi / 1.4
    assert abs(w * w - 9.81 * k * math.tanh(k * 0.45)) < 1e-10
    assert k == pytest.approx(2.52502, abs=5e-5)


def stokes2_fp64(x, y, t, a, k, w, h, x0, ramp_periods):
    """Textbook second-order Stokes surface and velocity with the quintic ramp."""
    q = min(max(t * w / (2 * math.pi * ramp_periods), 0.0), 1.0)
    a = a * q ** 3 * (10 - 15 * q + 6 * q * q)
    s = math.tanh(k * h)
    second = k * a * a * (3 - s * s) / (4 * s ** 3)
    ph = k * (x - x0) - w * t
    eta = a * math.cos(ph) + second * math.cos(2 * ph)
    slope = -k * a * math.sin(ph) - 2 * k * second * math.sin(2 * ph)
    sh = math.sinh(k * h)
    yy = y + h
    u = a * w / sh * math.cosh(k * yy) * math.cos(ph) + 0.75 * a * a * w * k / sh ** 4 * math.cosh(2 * k * yy) * math.cos(2 * ph)
    v = a * w / sh * math.sinh(k * yy) * math.sin(ph) + 0.75 * a * a * w * k / sh ** 4 * math.sinh(2 * k * yy) * math.sin(2 * ph)
    return eta, slope, u, v


def test_stokes_probes_known_answers():
    s = narrow_flume(n=(16, 12, 2))
    o = O.Oracle(s)
    o.op(86, 1)
    probe = o.get(O.FIELD_PROBE).reshape(8, 4)
    k, w = s.k, s.omega
    for i in range(8):
        t = s.ramp_periods * 2 * math.pi / w * (0.37 if i < 4 else 1)
        x = s.origin[0] + 0.17 + i * 0.23
        y = -s.depth + i * 0.07
        ref = stokes2_fp64(x, y, t, s.amplitude, k, w, s.depth, s.origin[0], s.ramp_periods)
        np.testing.assert_allclose(probe[i], ref, rtol=2e-5, atol=2e-6)


def small_flume(**kw):
    base = dict(n=(24, 48, 4), history_capacity=64, amplitude=0.0)
    base.update(kw)
    return narrow_flume(**base)


def wet_operator(o):
    """Dense A over wet cells from the face coefficients (work fields 6-12)."""
    nx, ny, nz = (int(v) for v in o.settings.n)
    N = o.N
    work = o.get(O.FIELD_WORK).reshape(14, N)
    geo = o.get(O.FIELD_GEO).reshape(N, 2)
    wet = geo[:, 1] > 0.5
    idx = -np.ones(N, np.int64)
    idx[wet] = np.arange(wet.sum())
    A = np.zeros((wet.sum(), wet.sum()))
    for i in np.flatnonzero(wet):
        x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
        face = 0
        for axis in range(3):
            for sign in (-1, 1):
                q = [x, y, z]
                q[axis] += sign
                q[2] %= nz
                if o.settings.periodic_x:
                    q[0] %= nx
                k = float(work[6 + face, i])
                A[idx[i], idx[i]] += k
                if 0 <= q[0] < nx and 0 <= q[1] < ny:
                    j = q[0] + nx * (q[1] + ny * q[2])
                    if wet[j]:
                        A[idx[i], idx[j]] -= k
                face += 1
    return A, wet, idx


def test_operator_is_spd_and_matches_spmv():
    s = small_flume(amplitude=0.06)
    o = O.Oracle(s)
    o.initial()
    A, wet, idx = wet_operator(o)
    assert np.allclose(A, A.T, rtol=0, atol=0), "masked operator must be exactly symmetric"
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > 0
    # op 5 with a random search direction (solving() requires pending, not converged)
    N = o.N
    c = o.control
    c[4], c[6], c[7] = 1, 0, 0
    o.set(O.FIELD_CONTROL, c)
    work = o.get(O.FIELD_WORK).reshape(14, N)
    p = RNG.normal(size=N).astype(np.float32)
    work[2] = p
    o.set(O.FIELD_WORK, work.ravel())
    o.op(5)
    Ap = o.get(O.FIELD_WORK).reshape(14, N)[3]
    ref = A @ p[wet].astype(np.float64)
    np.testing.assert_allclose(Ap[wet], ref, rtol=1e-5, atol=1e-3 * np.abs(ref).max())
    assert np.all(Ap[~wet] == 0)


def mg_pair_after_build(settings):
    o = O.Oracle(settings)
    o.initial()
    o.set_quota(1)
    # run one batch: the trial's setup builds the hierarchy from current coefficients
    o.batch(1)
    return o


def dense_transfer(o, l, active_c):
    """P (fine level l -> coarse l+1) from the oracle's own transfer rules,
    recomputed densely in FP64 from the level descriptors (trilinear with
    clamped ends, periodic in z and, with periodic_x, in x; aggregation below
    aggregate_from)."""
    levels = o.levels()
    nf = levels[l][:3]
    f = levels[l][5:8]
    agg = levels[l][15]
    nc = levels[l + 1][:3]
    Nf, Nc = np.prod(nf), np.prod(nc)
    P = np.zeros((Nf, Nc))
    for i in range(Nf):
        q = (i % nf[0], (i // nf[0]) % nf[1], i // (nf[0] * nf[1]))
        if agg:
            j = (q[0] // f[0]) + nc[0] * ((q[1] // f[1]) + nc[1] * (q[2] // f[2]))
            if active_c[j]:
                P[i, j] += 1
            continue
        axes = []
        for a in range(3):
            if f[a] == 1:
                axes.append([(q[a], 1.0)])
                continue
            pos = (q[a] + 0.5) / f[a] - 0.5
            b = math.floor(pos)
            t = pos - b
            opts = []
            for bit in (0, 1):
                nb = b + bit
                w = t if bit else 1 - t
                if a == 1:
                    if nb >= nc[1]:
                        continue
                    nb = max(nb, 0)
                elif a == 0:
                    nb = nb % nc[0] if o.settings.periodic_x else min(max(nb, 0), nc[0] - 1)
                else:
                    nb %= nc[2]
                opts.append((nb, w))
            axes.append(opts)
        for (x, wx), (y, wy), (z, wz) in itertools.product(*axes):
            j = x + nc[0] * (y + nc[1] * z)
            if active_c[j] and wx * wy * wz != 0:
                P[i, j] += wx * wy * wz
    return P


def level_matrix(o, l):
    """Dense matrix of level l >= 1 from the oracle's cached rows."""
    levels = o.levels()
    n = levels[l][:3]
    N = levels[l][8]
    width, row0 = levels[l][11], levels[l][12]
    rows = o.get(O.FIELD_MG_ROWS).reshape(-1, 2)
    cols = rows[:, 1].view(np.int32)
    A = np.zeros((N, N))
    for k in range(width):
        for i in range(N):
            e = row0 + k * N + i
            A[i, cols[e]] += rows[e, 0]
    return A


def test_galerkin_matches_dense_triple_product():
    s = small_flume(amplitude=0.06, n=(24, 48, 4))
    o = mg_pair_after_build(s)
    levels = o.levels()
    vec = o.get(O.FIELD_MG_VECTORS)
    act = []
    for l, m in enumerate(levels):
        off, cnt = m[3], m[8]
        act.append(vec[8 * off + 3 * cnt: 8 * off + 4 * cnt] > 0.5)
    # level 0 operator from the snapshot coefficients the preconditioner uses
    N = o.N
    snap = o.get(O.FIELD_SNAPSHOT).reshape(9, N)
    work = o.get(O.FIELD_WORK).reshape(14, N)
    work[6:13] = snap[0:7]
    o.set(O.FIELD_WORK, work.ravel())
    geo = o.get(O.FIELD_GEO).reshape(N, 2)
    geo[:, 1] = np.where(act[0], 1.0, np.where(geo[:, 1] < -0.5, -1.0, 0.0))
    o.set(O.FIELD_GEO, geo.ravel())
    A0, wet, idx = wet_operator(o)
    A = np.zeros((N, N))
    w = np.flatnonzero(wet)
    A[np.ix_(w, w)] = A0
    for l in range(len(levels) - 1):
        P = dense_transfer(o, l, act[l + 1])
        P[~act[l], :] = 0
        C = P.T @ A @ P
        Cl = level_matrix(o, l + 1)
        a = act[l + 1]
        err = np.abs(C[np.ix_(a, a)] - Cl[np.ix_(a, a)]).max() / np.abs(C[np.ix_(a, a)]).max()
        assert err < 2e-5, (l + 1, err)
        assert np.allclose(Cl[np.ix_(a, a)], Cl[np.ix_(a, a)].T, rtol=0, atol=0)
        A = np.where(np.outer(a, a), Cl, 0.0)


@pytest.mark.parametrize("periodic_x", [False, True])
def test_galerkin_entries_past_an_edge_are_zero(periodic_x):
    """A coarse row's stencil entry that points past a non-periodic face (where
    there is no coarse node) is exactly zero: P has no column there, so P^T A P
    has no entry.  The device sweeps rely on it to address every row by a
    linear delta with a single periodic correction."""
    from tests import parity as P
    s = small_flume(amplitude=0.06, n=(32, 48, 8), periodic_x=periodic_x)
    o = mg_pair_after_build(s)
    rows = o.get(O.FIELD_MG_ROWS).reshape(-1, 2)[:, 0]
    edges = P.coarse_edge_entries(o)
    assert len(edges) >= 2
    for r0, N, W, mask in edges:
        vals = rows[r0: r0 + W * N].reshape(W, N)
        assert mask.any()
        assert np.all(vals[mask] == 0)
        assert np.any(vals[~mask] != 0)


def test_coarse_cholesky_inverse_matches_numpy():
    s = small_flume(amplitude=0.06)
    o = mg_pair_after_build(s)
    L = len(o.levels())
    C = level_matrix(o, L - 1)
    N = C.shape[0]
    Linv = o.get(O.FIELD_MG_LINV).reshape(N, N)
    ref = np.linalg.inv(np.linalg.cholesky(C))
    np.testing.assert_allclose(Linv, ref, rtol=2e-4, atol=2e-4 * np.abs(ref).max())
    np.testing.assert_allclose(Linv.T @ Linv @ C, np.eye(N), atol=2e-4)


def test_vcycle_is_symmetric_positive():
    s = small_flume(amplitude=0.06)
    o = mg_pair_after_build(s)
    N = o.N
    geo = o.get(O.FIELD_GEO).reshape(N, 2)
    wet = geo[:, 1] > 0.5
    for _ in range(3):
        x = np.where(wet, RNG.normal(size=N), 0).astype(np.float32)
        y = np.where(wet, RNG.normal(size=N), 0).astype(np.float32)
        Mx, My = o.precondition(x), o.precondition(y)
        lhs, rhs = float(np.dot(Mx.astype(np.float64), y)), float(np.dot(x.astype(np.float64), My))
        assert abs(lhs - rhs) <= 1e-4 * max(abs(lhs), abs(rhs))
        assert float(np.dot(Mx.astype(np.float64), x)) > 0


def test_still_water_is_hydrostatic():
    s = small_flume(amplitude=0.0, perturbation=0.0)
    o = O.Oracle(s)
    o.initial()
    for _ in range(2):
        o.step()
    c = o.control
    assert c[2] == 2 and c[7] == 0
    bank = o.bank(int(c[0]))
    N = o.N
    faces = bank[N:N + o.faces]
    assert np.abs(faces).max() < 1e-4  # still water stays still
    assert c[16] < 1e-3 * 9810 * 0.45   # max |p - 9810 (eta - y)| over wet cells


def test_pressure_solve_matches_fp64_direct_solve():
    s = small_flume(amplitude=0.08, pressure_tolerance=1e-6)
    o = O.Oracle(s)
    o.initial()
    o.step()
    # replay the second step's setup batch and stop after the PCG converged
    o.set_quota(2)
    o.batch(2)
    c = o.control
    N = o.N
    work = o.get(O.FIELD_WORK).reshape(14, N)
    A, wet, idx = wet_operator(o)
    b = work[5][wet].astype(np.float64)
    cand = 1 - int(c[0]) if c[4] > 0.5 else int(c[0])
    p = o.bank(cand)[N + o.faces:N + o.faces + N]
    # projected p includes defect corrections; compare the residual of the system it solved
    p64 = spla.spsolve(sp.csr_matrix(A), b)
    rel = np.linalg.norm(A @ p[wet] - b) / np.linalg.norm(b)
    assert rel < 1e-3, rel
    assert np.linalg.norm(p[wet] - p64) / np.linalg.norm(p64) < 1e-3


def test_mass_ledger_closes_over_wave_steps():
    s = small_flume(amplitude=0.08, n=(40, 48, 4))
    o = O.Oracle(s)
    o.initial()
    for _ in range(4):
        o.step()
    c = o.control
    assert c[2] == 4 and c[7] == 0
    assert abs(c[28]) <= 1e-3 * c[26]
    F = o.bank(int(c[0]))[: o.N].astype(np.float64)
    assert F.min() >= 0 and F.max() <= 1
