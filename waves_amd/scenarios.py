"""Seeded inputs of the benchmark configurations (the reference has no RNG).

pressure_case(): the standalone pressure-solve input at a given grid -- the
reference beach's 1:6 solid bed, a seeded wavy free surface of mean depth
0.45 m and amplitude 0.08 m written as volume fractions the way the reference's
initial op writes its still surface (pressure_pass.glslinc:14-17), and face
velocities drawn from N(0, 0.1 m/s), zero on faces touching the bed or walls.
"""
from __future__ import annotations

import numpy as np

from .settings import Settings


def cell_centres(s: Settings):
    nx, ny, nz = (int(v) for v in s.n)
    hx, hy, hz = (np.float32(v) for v in s.spacing)
    ox, oy, oz = (np.float32(v) for v in s.origin)
    x = ox + (np.arange(nx, dtype=np.float32) + np.float32(0.5)) * hx
    y = oy + (np.arange(ny, dtype=np.float32) + np.float32(0.5)) * hy
    z = oz + (np.arange(nz, dtype=np.float32) + np.float32(0.5)) * hz
    return x, y, z


def solid_mask(s: Settings) -> np.ndarray:
    """pressure_core.glslinc:87-96 inside the grid: y <= -0.45 + max(0, x - 1.5)/slope.
    Shape (nz, ny, nx)."""
    x, y, _ = cell_centres(s)
    nz = int(s.n[2])
    if s.slope > 0:
        bed = np.float32(-0.45) + np.maximum(np.float32(0), x - np.float32(1.5)) / np.float32(s.slope)
        m = y[:, None] <= bed[None, :]
    else:
        m = np.zeros((len(y), len(x)), bool)
    return np.broadcast_to(m[None], (nz,) + m.shape).copy()


def wavy_surface(s: Settings, seed: int, amplitude: float = 0.08) -> np.ndarray:
    """eta(x, z): two seeded crossing cosine trains scaled to `amplitude`."""
    rng = np.random.default_rng(seed)
    x, _, z = cell_centres(s)
    X, Z = np.meshgrid(x.astype(np.float64), z.astype(np.float64))
    k1 = s.k * rng.uniform(0.8, 1.2)
    th = rng.uniform(-0.6, 0.6)
    k2 = 2 * np.pi / rng.uniform(0.6, 1.0)
    e = 0.6 * np.cos(k1 * X + rng.uniform(0, 2 * np.pi)) + 0.4 * np.cos(k2 * (X * np.cos(th) + Z * np.sin(th)) + rng.uniform(0, 2 * np.pi))
    e *= amplitude / np.abs(e).max()
    return e.astype(np.float32)  # (nz, nx)


def pressure_case(s: Settings, seed: int = 3, amplitude: float = 0.08, sigma: float = 0.1):
    """Returns a full bank (F, U, V, W, P=0) as float32 for wv_write_field."""
    nx, ny, nz = (int(v) for v in s.n)
    _, y, _ = cell_centres(s)
    eta = wavy_surface(s, seed, amplitude)
    solid = solid_mask(s)
    ph = y[None, :, None] - eta[:, None, :]                     # (nz, ny, nx)
    F = np.clip(np.float32(0.5) - ph / np.float32(s.spacing[1]), 0, 1).astype(np.float32)
    F[solid] = 0
    rng = np.random.default_rng(seed + 1000)
    NU, NV, NW = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    U = rng.normal(0, sigma, (nz, ny, nx + 1)).astype(np.float32)
    V = rng.normal(0, sigma, (nz, ny + 1, nx)).astype(np.float32)
    W = rng.normal(0, sigma, (nz + 1, ny, nx)).astype(np.float32)
    # faces touching a solid cell or the domain's solid sides carry no flow
    pad = np.ones((nz, ny + 2, nx + 2), bool)                  # outside = solid except top / ocean side
    pad[:, 1:-1, 1:-1] = solid
    pad[:, -1, :] = False                                       # open top
    pad[:, 1:-1, 0] = False                                     # ocean side (x < 0) is open
    U[(pad[:, 1:-1, :-1] | pad[:, 1:-1, 1:])] = 0
    V[(pad[:, :-1, 1:-1] | pad[:, 1:, 1:-1])] = 0
    sol = solid
    Wm = np.zeros((nz + 1, ny, nx), bool)
    Wm[1:-1] = sol[:-1] | sol[1:]
    Wm[0] = sol[0] | sol[-1]
    Wm[-1] = Wm[0]                                               # periodic duplicate face
    W[Wm] = 0
    W[-1] = W[0]
    return np.concatenate([F.ravel(), U.ravel(), V.ravel(), W.ravel(), np.zeros(nx * ny * nz, np.float32)])


def load_pressure_case(solver, s: Settings, seed: int = 3):
    """Write pressure_case() into both banks of a started solver and derive
    geometry, cell kinds and coefficients (ops 17, 18, 16).  A solve projects
    the candidate faces, so every standalone solve starts from a fresh load."""
    from . import solver as S
    bank = pressure_case(s, seed)
    solver.write(S.FIELD_BANK_A, bank)
    solver.write(S.FIELD_BANK_B, bank)
    for op in (17, 18, 16):
        solver.op(op)
    solver.synchronize()
