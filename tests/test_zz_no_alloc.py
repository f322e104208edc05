"""GPU: the allocation-free liquid volume (wv_liquid_volume_noalloc) against
wv_liquid_volume, the entry point it varies, bit for bit: two calls in a row,
the second on a smaller grid.  The caller's input and output buffers are
allocated here through the HIP runtime the library itself links (torch ships
a runtime of its own, which cannot share the device with it in one process);
the entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

from waves_amd import solver as S
from waves_amd.settings import narrow_flume

pytestmark = pytest.mark.gpu


def _hip():
    S.library()                                   # loads /opt/rocm's libamdhip64 with the solver library
    h = ctypes.CDLL("libamdhip64.so.7")           # the same runtime (same soname, already loaded)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _noalloc_volume(lib, hip, F: np.ndarray, spacing, stream=None):
    F = np.ascontiguousarray(F, np.float32)
    x, out = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(x), F.nbytes) == 0 and hip.hipMalloc(ctypes.byref(out), 8) == 0
    try:
        assert hip.hipMemcpy(x, F.ctypes.data, F.nbytes, H2D) == 0
        rc = lib.wv_liquid_volume_noalloc(x, ctypes.c_int64(F.size), ctypes.c_float(spacing[0]),
                                          ctypes.c_float(spacing[1]), ctypes.c_float(spacing[2]), out, stream)
        assert rc == 0, rc
        assert hip.hipDeviceSynchronize() == 0
        v = np.zeros(1, np.float64)
        assert hip.hipMemcpy(v.ctypes.data, out, 8, D2H) == 0
        return float(v[0])
    finally:
        hip.hipFree(x)
        hip.hipFree(out)


def test_noalloc_volume_matches_liquid_volume(gpu):
    lib = S.library()
    hip = _hip()
    cases = (narrow_flume(history_capacity=64), narrow_flume(n=(64, 48, 4), history_capacity=64))
    for k, s in enumerate(cases):          # the second call has the smaller input
        g = S.Solver(s)
        g.start()
        for _ in range(5 if k == 0 else 2):
            g.step()
        bank = int(g.control[0])
        ref = g.liquid_volume(bank)
        F = g.read(bank)[: s.cells]
        got = _noalloc_volume(lib, hip, F, s.spacing)
        print(f"call {k + 1}: {s.cells} cells, wv_liquid_volume {ref!r}, noalloc {got!r}", flush=True)
        assert np.float64(got).tobytes() == np.float64(ref).tobytes()
        g.close()


def test_noalloc_rejects_oversized_input(gpu):
    lib = S.library()
    rc = lib.wv_liquid_volume_noalloc(ctypes.c_void_p(16), ctypes.c_int64((1 << 24) + 1), ctypes.c_float(1),
                                      ctypes.c_float(1), ctypes.c_float(1), ctypes.c_void_p(16), None)
    assert rc == -1
