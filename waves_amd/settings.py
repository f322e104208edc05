"""Solver settings and push-constant packing.

Mirrors the settings dictionary of the reference's beach scene
(water/lab_runner.gd:19-26) and the 32-float parameter block that the
reference packs for every dispatch (water/pressure_gpu.gd:67-68,
water/coupled_gpu.gd:16-22, water/coastal_gpu.gd:21-25).  The same block is
handed to the HIP library through the C-ABI (``wv_settings.pc``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

#: push-constant indices (float32 slots), pressure_core.glslinc:34-43
PC_N = 0
PC_OP = 3
PC_H = 4
PC_DT_MAX = 7
PC_ORIGIN = 8
PC_ETA = 11
PC_PERIODIC_X = 12
PC_SLOPE = 13
PC_CORRUGATION = 14
PC_QUOTA = 15
PC_SHAPE = 16
PC_GRAVITY = 17
PC_MAX_FACE_CFL = 18
PC_HISTORY = 19
PC_WAVE = 20
PC_TOL = 24
PC_LINEAR_AIR = 25
PC_STOKES2 = 26
PC_FLAGS = 27
PC_FORCING = 28

#: numerics flag bits, water/coupled_gpu.gd:21
FLAG_HEIGHT_NORMALS = 1
FLAG_VALID_BAND = 2
FLAG_TRANSPORT_CACHE = 4
FLAG_INTERIOR = 8


def dispersion_k(period: float, depth: float, g: float = 9.81) -> float:
    """Wavenumber of omega^2 = g k tanh(k h) by 80 bisection steps on [0, 100]
    (water/lab_runner.gd:19-25, in double precision like GDScript floats)."""
    omega = 2.0 * math.pi / period
    low, high = 0.0, 100.0
    for _ in range(80):
        k = (low + high) * 0.5
        if g * k * math.tanh(k * depth) < omega * omega:
            low = k
        else:
            high = k
    return (low + high) * 0.5


@dataclass
class Settings:
    """The reference's settings dictionary (water/lab_runner.gd:26) plus the
    defaults injected by the solver classes' _init (water/pressure_gpu.gd:42-43,
    water/coupled_gpu.gd:9-15, water/coastal_gpu.gd:7-8)."""

    n: Sequence[int] = (310, 48, 80)
    spacing: Sequence[float] = (0.025, 0.025, 0.05)
    origin: Sequence[float] = (-1.75, -0.6, -2.0)
    eta: float = 0.0
    periodic_x: bool = False
    shape: int = 7
    amplitude: float = 0.08
    period: float = 1.4
    depth: float = 0.45
    zone_length: float = 2.5
    damping_rate: float = 12.0
    ramp_periods: float = 4.0
    stokes_second: bool = True
    history_capacity: int = 10000
    slope: float = 6.0
    corrugation: float = 0.0
    dt: float = 1.0 / 120.0
    gravity: float = 1.0
    max_face_cfl: bool = False
    pressure_tolerance: float = 1e-4
    linear_air: bool = False
    height_normals: bool = False
    active_advection: bool = True
    transport_cache: bool = True
    transport_interior: bool = False
    right_absorption: float = 1.0
    mg_smoothing: int = 3
    mg_lag: int = 8
    mg_aggregate_from: int = 2
    mg_maintenance: int = 2  # 0 rebuild, 1 structure, 2 lagged
    iterations_per_batch: int = 8
    perturbation: float = 0.0  # alongshore surface perturbation amplitude [m]
    seed: int = 0

    @property
    def k(self) -> float:
        return dispersion_k(self.period, self.depth)

    @property
    def omega(self) -> float:
        return 2.0 * math.pi / self.period

    def flags(self) -> int:
        return ((FLAG_HEIGHT_NORMALS if self.height_normals else 0)
                + (FLAG_VALID_BAND if self.active_advection else 0)
                + (FLAG_TRANSPORT_CACHE if self.transport_cache else 0)
                + (FLAG_INTERIOR if self.transport_interior else 0))

    def push_constants(self, op: int = 0, quota: float = 0.0) -> np.ndarray:
        """The 32-float block of water/pressure_gpu.gd:68 + coupled_gpu.gd:18-21
        + coastal_gpu.gd:23-24, as float32."""
        pc = np.zeros(32, np.float32)
        pc[0:3] = self.n
        pc[PC_OP] = op
        pc[4:7] = self.spacing
        pc[PC_DT_MAX] = self.dt
        pc[8:11] = self.origin
        pc[PC_ETA] = self.eta
        pc[PC_PERIODIC_X] = 1 if self.periodic_x else 0
        pc[PC_SLOPE] = self.slope
        pc[PC_CORRUGATION] = self.corrugation
        pc[PC_QUOTA] = quota
        pc[PC_SHAPE] = self.shape
        pc[PC_GRAVITY] = self.gravity
        pc[PC_MAX_FACE_CFL] = 1 if self.max_face_cfl else 0
        pc[PC_HISTORY] = self.history_capacity
        pc[20:24] = (self.amplitude, self.k, self.omega, self.depth)
        pc[PC_TOL] = self.pressure_tolerance
        pc[PC_LINEAR_AIR] = 1 if self.linear_air else 0
        pc[PC_STOKES2] = 1 if self.stokes_second else 0
        pc[PC_FLAGS] = self.flags()
        pc[28:32] = (self.origin[0], self.zone_length, self.damping_rate, self.ramp_periods)
        return pc

    def options(self) -> np.ndarray:
        """Eight option floats shared by the oracle and the HIP library:
        right absorption, MG smoothing / lag / aggregate_from / maintenance,
        PCG iterations per batch, perturbation amplitude and seed."""
        return np.array([self.right_absorption, self.mg_smoothing, self.mg_lag,
                         self.mg_aggregate_from, self.mg_maintenance,
                         self.iterations_per_batch, self.perturbation, self.seed],
                        np.float32)

    @property
    def cells(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])


def default_beach(**overrides) -> Settings:
    """The reference beach (water/lab_runner.gd:26): 310x48x80, 1:6 slope."""
    return Settings(**overrides)


def benchmark_beach(**overrides) -> Settings:
    """Default beach (the plain method at the reference's own PCG tolerance,
    1e-4, water/coupled_gpu.gd:21) with the seeded 1e-3 m alongshore
    perturbation: the 20 sim-s benchmark configuration at 310x48x80.  Six PCG
    iteration slots per submitted batch instead of eight: the device state
    machine gates every op, so results are bitwise identical for any split.
    In the breaking phase (5.5 PCG iterations per step) six slots finish most
    trials in one batch without replaying many idle slots: 7.14 ms per step
    against 7.33 with four slots and 7.40 with eight."""
    base = dict(pressure_tolerance=1e-4, perturbation=1e-3, seed=1, iterations_per_batch=6)
    base.update(overrides)
    return Settings(**base)


def narrow_flume(**overrides) -> Settings:
    """124x48x8 flume (3.1 x 1.2 x 0.4 m).  The bed toe is hard-coded at world
    x = 1.5 m (pressure_core.glslinc:94), so origin.x = 0.5 puts the toe 1.0 m
    from the inflow; the relaxation zone is shortened to 0.75 m so that it ends
    before the toe."""
    base = dict(n=(124, 48, 8), origin=(0.5, -0.6, -0.2), zone_length=0.75,
                pressure_tolerance=1e-6, perturbation=1e-3, seed=1,
                history_capacity=2400)
    base.update(overrides)
    return Settings(**base)
