"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WAVES_ORACLE_LIB: a differently built oracle library (the mutation sweep of
# tests/oracle_mutations.py loads deliberately broken builds through it)
LIB = os.environ.get("WAVES_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

FIELD_BANK_A, FIELD_BANK_B, FIELD_WORK, FIELD_GEO, FIELD_CONTROL = 0, 1, 2, 3, 4
FIELD_PLANE, FIELD_MOTION, FIELD_IMG_PLANES, FIELD_IMG_KINDS = 5, 6, 7, 8
FIELD_HISTORY, FIELD_COASTAL_HISTORY, FIELD_PROBE, FIELD_PARTIAL = 9, 10, 11, 12
FIELD_MG_VECTORS, FIELD_SNAPSHOT, FIELD_AUDIT, FIELD_MG_ROWS, FIELD_MG_LINV = 13, 14, 15, 16, 17
FIELD_PREVIOUS = 19

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i64, f32p = ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_float)
        L.wo_create.restype = vp
        L.wo_create.argtypes = [f32p, f32p]
        L.wo_destroy.argtypes = [vp]
        L.wo_field_count.restype = i64
        L.wo_field_count.argtypes = [vp, ctypes.c_int]
        L.wo_get.argtypes = [vp, ctypes.c_int, vp, i64]
        L.wo_set.argtypes = [vp, ctypes.c_int, vp, i64]
        L.wo_op.argtypes = [vp, ctypes.c_int, i64]
        L.wo_mg_record.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.wo_initial.argtypes = [vp]
        L.wo_set_quota.argtypes = [vp, ctypes.c_float]
        L.wo_set_scatter_order.argtypes = [vp, ctypes.c_int]
        L.wo_batch.argtypes = [vp, ctypes.c_float]
        L.wo_levels.argtypes = [vp]
        L.wo_level_meta.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.wo_precondition.argtypes = [vp, f32p, f32p]
        L.wo_cube_volume.restype = ctypes.c_float
        L.wo_cube_volume.argtypes = [f32p, ctypes.c_float]
        L.wo_inverse_volume.restype = ctypes.c_float
        L.wo_inverse_volume.argtypes = [f32p, ctypes.c_float]
        _lib = L
    return _lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Oracle:
    """One CPU restatement of the solver, built from a waves_amd Settings."""

    def __init__(self, settings):
        self.settings = settings
        self.pc = settings.push_constants()
        self.opts = settings.options()
        self.h = lib().wo_create(_fp(self.pc), _fp(self.opts))
        if not self.h:
            raise RuntimeError("wo_create failed")
        nx, ny, nz = (int(v) for v in settings.n)
        self.N = nx * ny * nz
        self.NU, self.NV, self.NW = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
        self.faces = self.NU + self.NV + self.NW
        self.stride = 2 * self.N + self.faces

    def close(self):
        if self.h:
            lib().wo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, field: int, dtype=np.float32) -> np.ndarray:
        n = lib().wo_field_count(self.h, field)
        out = np.empty(n, dtype)
        if lib().wo_get(self.h, field, out.ctypes.data, n) != 0:
            raise RuntimeError("wo_get failed")
        return out

    def set(self, field: int, values: np.ndarray):
        v = np.ascontiguousarray(values)
        if lib().wo_set(self.h, field, v.ctypes.data, v.size) != 0:
            raise RuntimeError("wo_set failed")

    def op(self, op: int, groups: int | None = None):
        if groups is None:
            groups = (self.stride + 127) // 128
        lib().wo_op(self.h, op, groups)

    def set_quota(self, quota: float):
        lib().wo_set_quota(self.h, quota)

    def set_scatter_order(self, reverse: bool):
        """Galerkin scatter additions in descending (True) or ascending entry order."""
        lib().wo_set_scatter_order(self.h, int(bool(reverse)))

    def initial(self):
        lib().wo_initial(self.h)

    def batch(self, quota: float):
        lib().wo_batch(self.h, quota)

    def mg_record(self, build: bool, initial: bool = False):
        lib().wo_mg_record(self.h, int(build), int(initial))

    @property
    def control(self) -> np.ndarray:
        return self.get(FIELD_CONTROL)

    def bank(self, which: int) -> np.ndarray:
        return self.get(FIELD_BANK_A if which == 0 else FIELD_BANK_B)

    def levels(self):
        L = lib().wo_levels(self.h)
        out = []
        for l in range(L):
            m = (ctypes.c_int * 16)()
            lib().wo_level_meta(self.h, l, m)
            out.append(list(m))
        return out

    def precondition(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float32)
        z = np.empty_like(r)
        lib().wo_precondition(self.h, _fp(r), _fp(z))
        return z

    def step(self, max_batches: int = 64) -> int:
        """Run batches until one more step is accepted or the run stops.
        Returns the number of batches used."""
        target = self.control[2] + 1
        for b in range(1, max_batches + 1):
            self.batch(float(target))
            c = self.control
            f = c[7]
            if c[2] >= target or (f > 0.5 and not ((f == 20 or 10 <= f <= 12) and c[29] < 2)):
                return b
        return max_batches


def cube_volume(normal, alpha: float) -> float:
    n = np.ascontiguousarray(normal, np.float32)
    return float(lib().wo_cube_volume(_fp(n), ctypes.c_float(alpha)))


def inverse_volume(normal, fraction: float) -> float:
    n = np.ascontiguousarray(normal, np.float32)
    return float(lib().wo_inverse_volume(_fp(n), ctypes.c_float(fraction)))
