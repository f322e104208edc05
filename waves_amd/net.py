"""Learned warm start: the U-Net of include/waves_net.h over the C-ABI.

No reference counterpart.  Training data are (F, r, p_prev, wet) inputs and
normalised pressure-correction targets built from solver runs
(waves_amd/learned.py); the network runs and trains only on the device
(libwaves_hip.so), there is no host fallback.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from .solver import WavesError, library

IN_CHANNELS = 8


def level_dims(n: Sequence[int], levels: int = 4):
    dims = [tuple(int(v) for v in n)]
    for _ in range(levels - 1):
        dims.append(tuple((v + 1) // 2 for v in dims[-1]))
    return dims


def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    """float -> bfloat16 bit patterns (round to nearest even), uint16."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[np.isnan(f)] = 0x7FC0
    return r


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Net:
    """A network for grids n = (nx, ny, nz) and `batch` samples per pass."""

    def __init__(self, n: Sequence[int], batch: int = 1):
        self._lib = library()
        self.n = tuple(int(v) for v in n)
        self.batch = int(batch)
        self.cells = self.n[0] * self.n[1] * self.n[2]
        d2 = level_dims(self.n)[2]
        self.cells2 = d2[0] * d2[1] * d2[2]
        self._h = ctypes.c_void_p()
        rc = self._lib.wv_net_create(self.n[0], self.n[1], self.n[2], self.batch, ctypes.byref(self._h))
        if rc != 0:
            raise WavesError(f"wv_net_create failed ({rc})")
        self.param_count = int(self._lib.wv_net_param_count())

    def close(self):
        if self._h:
            self._lib.wv_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            msg = self._lib.wv_net_last_error(self._h) if self._h else b""
            raise WavesError(f"waves net error {rc}: {msg.decode(errors='replace') if msg else ''}")

    @staticmethod
    def _f(a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    @staticmethod
    def _u(a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count:
            raise ValueError("parameter count")
        self._check(self._lib.wv_net_set_params(self._h, self._f(p), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.param_count, np.float32)
        self._check(self._lib.wv_net_get_params(self._h, self._f(p), p.size))
        return p

    def forward(self, x_bits: np.ndarray):
        """x_bits: (batch*cells*8) bf16 bits -> (qf (batch, cells), qc (batch, cells2))."""
        x = np.ascontiguousarray(x_bits, np.uint16).ravel()
        assert x.size == self.batch * self.cells * IN_CHANNELS
        qf = np.empty(self.batch * self.cells, np.float32)
        qc = np.empty(self.batch * self.cells2, np.float32)
        self._check(self._lib.wv_net_forward(self._h, self._u(x), self._f(qf), self._f(qc)))
        return qf.reshape(self.batch, -1), qc.reshape(self.batch, -1)

    def loss_grad(self, x_bits: np.ndarray, target: np.ndarray):
        x = np.ascontiguousarray(x_bits, np.uint16).ravel()
        t = np.ascontiguousarray(target, np.float32).ravel()
        assert x.size == self.batch * self.cells * IN_CHANNELS and t.size == self.batch * self.cells
        loss = np.zeros(1, np.float32)
        g = np.empty(self.param_count, np.float32)
        self._check(self._lib.wv_net_loss_grad(self._h, self._u(x), self._f(t), self._f(loss), self._f(g)))
        return float(loss[0]), g

    def set_dataset(self, x_bits: np.ndarray, target: np.ndarray):
        x = np.ascontiguousarray(x_bits, np.uint16)
        t = np.ascontiguousarray(target, np.float32)
        count = t.size // self.cells
        assert x.size == count * self.cells * IN_CHANNELS
        self._check(self._lib.wv_net_set_dataset(self._h, self._u(x), self._f(t), count))
        self.data_count = count

    def train(self, steps: int, train_count: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.empty(int(steps), np.float32)
        self._check(self._lib.wv_net_train(self._h, int(steps), int(train_count), float(lr),
                                           ctypes.c_uint64(int(seed)), self._f(losses)))
        return losses

    def layers(self):
        """[dict(level, cells, coarse_cells, ca, cb, cout)] of every layer in
        forward order: ca input channels at the layer's level, cb upsampled
        from the next level (coarse_cells), cout outputs."""
        dims = level_dims(self.n) + [(0, 0, 0)]
        out = []
        for l in range(int(self._lib.wv_net_layer_count())):
            v = [ctypes.c_int() for _ in range(4)]
            self._check(self._lib.wv_net_layer_info(l, *[ctypes.byref(x) for x in v]))
            lv, ca, cb, co = (x.value for x in v)
            d, dc = dims[lv], dims[lv + 1]
            out.append({"level": lv, "cells": d[0] * d[1] * d[2] * self.batch,
                        "coarse_cells": dc[0] * dc[1] * dc[2] * self.batch, "ca": ca, "cb": cb, "cout": co})
        return out

    def time_layers(self, reps: int = 20) -> np.ndarray:
        """Mean device microseconds per launch of every layer's convolution
        (inputs from one forward pass of the batch in the input buffer)."""
        n = int(self._lib.wv_net_layer_count())
        us = np.zeros(n, np.float32)
        self._check(self._lib.wv_net_time_layers(self._h, int(reps), self._f(us), n))
        return us

    def eval(self, first: int, count: int) -> float:
        loss = np.zeros(1, np.float32)
        self._check(self._lib.wv_net_eval(self._h, int(first), int(count), self._f(loss)))
        return float(loss[0])
