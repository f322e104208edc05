"""Host-side mirror of the reference solver runtime over the HIP C-ABI.

The reference's host is GDScript: water/pressure_gpu.gd (lifecycle, request /
request_time quota, telemetry decode), water/coupled_gpu.gd and
water/coastal_gpu.gd (extra telemetry), water/work_scheduler.gd (frame-driven
submission with one asynchronous control readback).  This module keeps those
names and meanings on top of libwaves_hip.so (include/waves.h) through ctypes.
There is no fallback: without the built library every entry raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Dict, Optional

import numpy as np

from .settings import Settings

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwaves_hip.so")

FIELD_BANK_A, FIELD_BANK_B, FIELD_WORK, FIELD_GEO, FIELD_CONTROL = 0, 1, 2, 3, 4
FIELD_PLANE, FIELD_MOTION, FIELD_IMG_PLANES, FIELD_IMG_KINDS = 5, 6, 7, 8
FIELD_HISTORY, FIELD_COASTAL_HISTORY, FIELD_PROBE, FIELD_PARTIAL = 9, 10, 11, 12
FIELD_MG_VECTORS, FIELD_SNAPSHOT, FIELD_AUDIT, FIELD_MG_ROWS, FIELD_MG_LINV = 13, 14, 15, 16, 17
FIELD_PRESSURE64, FIELD_PREVIOUS, FIELD_STEP_TIMES, FIELD_QUEUES = 18, 19, 20, 21
OP_SPMV_DOT = 105  # ops 5, 6, 7 in the batch's one-pass form (no reference counterpart)
MAX_INFLIGHT = 3

#: every entry point declared in include/waves.h
EXPORTS = ("wv_create", "wv_destroy", "wv_reset", "wv_request_steps", "wv_request_time",
           "wv_advance", "wv_poll_control", "wv_field_bytes", "wv_read_field",
           "wv_write_field", "wv_op", "wv_mg_record", "wv_synchronize", "wv_last_error",
           "wv_levels", "wv_plan_levels", "wv_time_kernel", "wv_pressure_solve", "wv_set_precision",
           "wv_liquid_volume", "wv_get_schedule_stats", "wv_liquid_volume_noalloc", "wv_save_state",
           "wv_load_state")
#: every entry point declared in include/waves_net.h
NET_EXPORTS = ("wv_net_param_count", "wv_net_create", "wv_net_destroy", "wv_net_set_params",
               "wv_net_get_params", "wv_net_forward", "wv_net_loss_grad", "wv_net_set_dataset",
               "wv_net_train", "wv_net_eval", "wv_net_last_error", "wv_solver_attach_net",
               "wv_net_layer_count", "wv_net_layer_info", "wv_net_time_layers")
SOLVE_ZERO_MEAN, SOLVE_GIVEN_RHS = 1, 2


class WavesError(RuntimeError):
    pass


def stopped(c: np.ndarray) -> bool:
    """True when the device has stopped on a failure: a non-zero c[7] that is
    not one of the recoverable codes still within its retry budget (CFL 20,
    VOF bounds 10-12, at most 2 retries; coupled_motion.glslinc:351-354,
    water/coupled_gpu.gd:61-62)."""
    f = c[7]
    if f < 0.5:
        return False
    recoverable = (f == 20 or 10 <= f <= 12) and c[29] < 2
    return not recoverable


class _Settings(ctypes.Structure):
    _fields_ = [("pc", ctypes.c_float * 32), ("opts", ctypes.c_float * 8)]


class _ScheduleStats(ctypes.Structure):
    _fields_ = [("submissions", ctypes.c_int64), ("readbacks", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("inflight", ctypes.c_int), ("max_inflight", ctypes.c_int), ("inflight_batches", ctypes.c_int64),
                ("gpu_ms", ctypes.c_double), ("record_us", ctypes.c_double), ("ema_batch_ms", ctypes.c_double),
                ("accepted_time", ctypes.c_double), ("accepted_steps", ctypes.c_double), ("r_fluid", ctypes.c_double)]


class _SolveReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("failure", ctypes.c_int),
                ("restarts", ctypes.c_int), ("device_ms", ctypes.c_float), ("wall_ms", ctypes.c_float),
                ("true_rel_residual", ctypes.c_double), ("recurrence_rel_residual", ctypes.c_float)]


_lib = None
_lib_lock = threading.Lock()


def library() -> ctypes.CDLL:
    """Load libwaves_hip.so (raises if it has not been built)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise WavesError(f"{LIB_PATH} is missing: run `python -m waves_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        fp = ctypes.POINTER(ctypes.c_float)
        L.wv_create.argtypes = [ctypes.POINTER(_Settings), ctypes.POINTER(vp)]
        L.wv_destroy.argtypes = [vp]
        L.wv_reset.argtypes = [vp]
        L.wv_request_steps.argtypes = [vp, i64]
        L.wv_request_time.argtypes = [vp, ctypes.c_double]
        L.wv_advance.argtypes = [vp, i32, ctypes.c_float]
        L.wv_get_schedule_stats.argtypes = [vp, ctypes.POINTER(_ScheduleStats)]
        L.wv_save_state.argtypes = [vp, ctypes.c_char_p]
        L.wv_load_state.argtypes = [vp, ctypes.c_char_p]
        L.wv_liquid_volume_noalloc.argtypes = [vp, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp]
        L.wv_poll_control.argtypes = [vp, fp, ctypes.POINTER(i32), i32]
        L.wv_field_bytes.restype = i64
        L.wv_field_bytes.argtypes = [vp, i32]
        L.wv_read_field.argtypes = [vp, i32, vp, i64]
        L.wv_write_field.argtypes = [vp, i32, vp, i64]
        L.wv_op.argtypes = [vp, i32]
        L.wv_mg_record.argtypes = [vp, i32, i32]
        L.wv_synchronize.argtypes = [vp]
        L.wv_last_error.restype = ctypes.c_char_p
        L.wv_last_error.argtypes = [vp]
        L.wv_levels.argtypes = [vp, ctypes.POINTER(i32), i32]
        L.wv_plan_levels.argtypes = [ctypes.POINTER(_Settings), ctypes.POINTER(i32), i32]
        L.wv_time_kernel.argtypes = [vp, i32, i32, fp]
        L.wv_pressure_solve.argtypes = [vp, i32, ctypes.c_float, i32, i32, ctypes.POINTER(_SolveReport)]
        L.wv_set_precision.argtypes = [vp, i32]
        L.wv_liquid_volume.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double)]
        u16p = ctypes.POINTER(ctypes.c_uint16)
        L.wv_net_param_count.restype = i64
        L.wv_net_param_count.argtypes = []
        L.wv_net_create.argtypes = [i32, i32, i32, i32, ctypes.POINTER(vp)]
        L.wv_net_destroy.argtypes = [vp]
        L.wv_net_destroy.restype = None
        L.wv_net_set_params.argtypes = [vp, fp, i64]
        L.wv_net_get_params.argtypes = [vp, fp, i64]
        L.wv_net_forward.argtypes = [vp, u16p, fp, fp]
        L.wv_net_layer_count.argtypes = []
        L.wv_net_layer_info.argtypes = [i32] + [ctypes.POINTER(i32)] * 4
        L.wv_net_time_layers.argtypes = [vp, i32, fp, i32]
        L.wv_net_loss_grad.argtypes = [vp, u16p, fp, fp, fp]
        L.wv_net_set_dataset.argtypes = [vp, u16p, fp, i32]
        L.wv_net_train.argtypes = [vp, i32, i32, ctypes.c_float, ctypes.c_uint64, fp]
        L.wv_net_eval.argtypes = [vp, i32, i32, fp]
        L.wv_net_last_error.restype = ctypes.c_char_p
        L.wv_net_last_error.argtypes = [vp]
        L.wv_solver_attach_net.argtypes = [vp, vp]
        _lib = L
        return L


def _pack(settings: Settings) -> _Settings:
    s = _Settings()
    s.pc[:] = [float(v) for v in settings.push_constants()]
    s.opts[:] = [float(v) for v in settings.options()]
    return s


def plan_levels(settings: Settings):
    """The multigrid level plan (16 ints per level, the reference's `meta`,
    water/pressure_multigrid.gd:33-69) computed on the host, no device needed."""
    meta = (ctypes.c_int * (16 * 12))()
    L = library().wv_plan_levels(ctypes.byref(_pack(settings)), meta, 12)
    if L < 0:
        raise WavesError(f"wv_plan_levels failed ({L})")
    return [list(meta[16 * l:16 * l + 16]) for l in range(L)]


class Solver:
    """The reference's coastal solver object (water/presentation/coastal_view.gd
    -> coastal_gpu.gd -> coupled_gpu.gd -> pressure_gpu.gd) on the HIP library."""

    def __init__(self, settings: Settings):
        self.settings = settings
        self.config = settings
        self.closed = False
        self.ready = False
        self.last_stats: Dict = {}
        self._h = ctypes.c_void_p()
        s = _pack(settings)
        self._lib = library()
        rc = self._lib.wv_create(ctypes.byref(s), ctypes.byref(self._h))
        if rc != 0:
            raise WavesError(f"wv_create failed ({rc})")

    # -- lifecycle: water/pressure_gpu.gd:45-59 ------------------------------
    def start(self):
        """Run the initial command list (water/coupled_gpu.gd:34-37)."""
        self._check(self._lib.wv_reset(self._h))
        self.ready = True

    def close(self):
        if self.closed:
            return
        self.closed = True
        if self._h:
            self._lib.wv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            msg = self._lib.wv_last_error(self._h) if self._h else b""
            raise WavesError(f"waves_hip error {rc}: {msg.decode(errors='replace') if msg else ''}")

    # -- quota: water/pressure_gpu.gd:48-54 ----------------------------------
    def request(self, steps: int):
        self._check(self._lib.wv_request_steps(self._h, int(steps)))

    def request_time(self, seconds: float):
        assert np.isfinite(seconds) and seconds >= 0
        self._check(self._lib.wv_request_time(self._h, float(seconds)))

    # -- scheduling: water/work_scheduler.gd:41-87 ---------------------------
    def frame(self, max_batches: int = 1, budget_ms: float = 0.0) -> int:
        """One submission of up to `max_batches` batches (as many as fit in
        `budget_ms` when > 0) and one control readback, non-blocking; returns
        the batches enqueued (0 while MAX_INFLIGHT submissions are unfinished)."""
        n = self._lib.wv_advance(self._h, int(max_batches), float(budget_ms))
        if n < 0:
            self._check(n)
        return n

    def poll(self, wait=False) -> Optional[np.ndarray]:
        """Newest control readback; None if nothing new arrived (and not
        waiting).  wait: False/0 no blocking, True/1 until all enqueued work is
        done, 2 until the oldest unfinished submission is done."""
        out = np.zeros(96, np.float32)
        fresh = ctypes.c_int(0)
        self._check(self._lib.wv_poll_control(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(fresh), int(wait)))
        if fresh.value or wait:
            self.last_stats = self._stats(out)
            return out
        return None

    def schedule_stats(self) -> Dict:
        """Scheduler telemetry: submissions, readbacks, gpu_ms, record_us,
        R_fluid (water/work_scheduler.gd:18-40)."""
        st = _ScheduleStats()
        self._check(self._lib.wv_get_schedule_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _ScheduleStats._fields_}

    def step_times(self) -> np.ndarray:
        """(accepted steps, 4) uint64: pressure ns of the accepted trial, of all
        trials since the previous acceptance, trial count, sections timed."""
        a = self.read(FIELD_STEP_TIMES, np.uint64).reshape(-1, 4)
        return a[: int(self.control[2])]

    def run_steps(self, count: int, max_batches: int = 8, budget_ms: float = 0.0):
        """Advance `count` more accepted steps with up to MAX_INFLIGHT
        submissions in flight, sizing each submission by the observed batches
        per accepted step so that few batches run past the quota; returns the
        control vector when done (or stopped)."""
        c = self.poll(wait=True)
        target = int(c[2]) + int(count)
        self.request(target)
        st = self.schedule_stats()
        b0, s0 = st["batches"], int(c[2])
        per_step = 2.0
        while True:
            c = self.poll(wait=False)
            c = self._last_control() if c is None else c
            st = self.schedule_stats()
            done = int(c[2])
            if done >= target or stopped(c):
                if st["inflight"] == 0:
                    return c
                self.poll(wait=2)
                continue
            if done > s0:
                per_step = max(1.0, (st["batches"] - st["inflight_batches"] - b0) / (done - s0))
            need = int(np.ceil((target - done) * per_step)) - st["inflight_batches"]
            if need > 0 and st["inflight"] < MAX_INFLIGHT:
                self.frame(min(max_batches, need), budget_ms)
            else:
                self.poll(wait=2)

    def _last_control(self) -> np.ndarray:
        out = np.zeros(96, np.float32)
        fresh = ctypes.c_int(0)
        self._check(self._lib.wv_poll_control(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(fresh), 0))
        return out

    def synchronize(self):
        self._check(self._lib.wv_synchronize(self._h))

    # -- telemetry: water/pressure_gpu.gd:133-137, coupled_gpu.gd:58-65,
    #    coastal_gpu.gd:34-37 -------------------------------------------------
    @staticmethod
    def _stats(v: np.ndarray) -> Dict:
        reported = v.copy()
        recoverable = (v[7] == 20 or 10 <= v[7] <= 12) and v[29] < 2
        if recoverable:
            reported[7] = 0
        r = reported
        p = {"pressure_backend": "multigrid", "accepted_index": int(r[0]), "accepted_time": float(r[1]),
             "accepted_steps": int(r[2]), "rejections": int(r[3]), "pending": bool(r[4] > 0.5),
             "converged": bool(r[6] > 0.5), "failure": int(r[7]), "iterations": int(r[8]),
             "relative_residual": float(r[13]), "max_dt_divergence": float(r[14]),
             "max_face_speed": float(r[15]), "hydrostatic_error_pa": float(r[16]),
             "volume_m3": float(r[17]), "total_iterations": int(r[18])}
        p.update({"rhs_integrated_divergence_l2": float(v[62]), "residual_integrated_divergence_l2": float(v[63]),
                  "true_relative_residual": float(v[60]), "pressure_defect_corrections": int(v[58]),
                  "boundary_outflow_m3": float(v[40]), "tentative_outflow_m3": float(v[39]),
                  "raw_compression_volume_m3": float(v[38]), "compression_volume_m3": float(v[35]),
                  "compression_coverage_errors": int(v[36]), "step_start_volume_m3": float(v[37]),
                  "trial_count": int(v[32]), "trial_failure": int(v[7]), "retry_count": int(v[29]),
                  "last_rejection_code": int(v[33]), "last_rejected_dt": float(v[34]), "dt": float(v[20]),
                  "signed_roundoff_m3": float(v[24]), "absolute_roundoff_m3": float(v[25]),
                  "max_CFL": float(v[27]), "mass_defect_m3": float(v[28]), "initial_volume_m3": float(v[26]),
                  "coupled": True})
        p.update({"forcing_min_F": float(v[88]), "forcing_max_F": float(v[89]), "forcing_volume_m3": float(v[72]),
                  "ocean_inflow_m3": float(v[73]), "ocean_outflow_m3": float(v[74]),
                  "forcing_momentum_kg_mps": [float(x) for x in v[75:78]],
                  "staged_forcing_volume_m3": float(v[64]), "staged_ocean_inflow_m3": float(v[65]),
                  "staged_ocean_outflow_m3": float(v[66]), "gauges_m": [float(x) for x in v[80:88]]})
        return p

    # -- parity / export I/O -------------------------------------------------
    def field_bytes(self, field: int) -> int:
        return int(self._lib.wv_field_bytes(self._h, field))

    def read(self, field: int, dtype=np.float32) -> np.ndarray:
        n = self.field_bytes(field)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype)
        self._check(self._lib.wv_read_field(self._h, field, out.ctypes.data, n))
        return out

    def write(self, field: int, values: np.ndarray):
        v = np.ascontiguousarray(values)
        self._check(self._lib.wv_write_field(self._h, field, v.ctypes.data, v.nbytes))

    def save_state(self, path: str):
        """Write the whole solver state to a state file (wv_save_state)."""
        self._check(self._lib.wv_save_state(self._h, os.fsencode(path)))

    def load_state(self, path: str):
        """Load a state file written for the same settings (wv_load_state)."""
        self._check(self._lib.wv_load_state(self._h, os.fsencode(path)))

    def op(self, op: int):
        self._check(self._lib.wv_op(self._h, int(op)))

    def mg_record(self, build: bool, initial: bool = False):
        self._check(self._lib.wv_mg_record(self._h, int(build), int(initial)))

    @property
    def control(self) -> np.ndarray:
        return self.read(FIELD_CONTROL)

    def levels(self):
        meta = (ctypes.c_int * (16 * 12))()
        L = self._lib.wv_levels(self._h, meta, 12)
        return [list(meta[16 * l:16 * l + 16]) for l in range(L)]

    def time_kernel(self, kernel: int, reps: int = 50) -> float:
        out = ctypes.c_float(0)
        self._check(self._lib.wv_time_kernel(self._h, kernel, reps, ctypes.byref(out)))
        return float(out.value)

    def pressure_solve(self, mode: int = 0, tolerance: float = 1e-6, max_iterations: int = 4096,
                       zero_mean: bool = True, given_rhs: bool = False) -> Dict:
        """Standalone pressure solve on the current state (wv_pressure_solve):
        mode 0 plain FP32 MGPCG, mode 1 BF16 V-cycle / FP64 dots and true
        residual, mode 2 mode 1 plus the attached network's warm start.
        given_rhs: b is work field 5 and the start is the candidate pressure."""
        r = _SolveReport()
        flags = (SOLVE_GIVEN_RHS if given_rhs else (SOLVE_ZERO_MEAN if zero_mean else 0))
        self._check(self._lib.wv_pressure_solve(self._h, int(mode), float(tolerance), int(max_iterations),
                                                flags, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _SolveReport._fields_}

    def liquid_volume(self, bank: Optional[int] = None) -> float:
        """FP64 liquid volume of a bank (default: the accepted one)."""
        if bank is None:
            bank = int(self.control[0])
        v = ctypes.c_double()
        self._check(self._lib.wv_liquid_volume(self._h, int(bank), ctypes.byref(v)))
        return v.value

    def set_precision(self, mode: int) -> None:
        """Batch program precision: 0 the reference's FP32 PCG, 1 BF16 V-cycle
        operators with FP64 dot products (wv_set_precision)."""
        self._check(self._lib.wv_set_precision(self._h, int(mode)))

    def attach_net(self, net) -> None:
        """Use `net` (waves_amd.net.Net on this grid, batch 1) for mode 2."""
        self._check(self._lib.wv_solver_attach_net(self._h, net._h if net is not None else None))
        self._net = net

    def step(self, max_batches: int = 64) -> int:
        """Advance until one more step is accepted (or the run stops); one
        control readback per batch. Returns the batches used."""
        target = int(self.control[2]) + 1
        self.request(target)
        for b in range(1, max_batches + 1):
            self.frame(1)
            c = self.poll(wait=True)
            if c[2] >= target or stopped(c):
                return b
        return max_batches
