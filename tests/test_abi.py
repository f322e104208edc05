"""The C-ABI library loads without a GPU and exports every entry point that
include/waves.h declares; the host-side settings packing matches the
reference's parameter block (water/pressure_gpu.gd:68, coupled_gpu.gd:18-21,
coastal_gpu.gd:23-24).  No compute calls here."""
import ctypes
import os
import re

import numpy as np

from waves_amd import solver as S
from waves_amd.settings import Settings, default_beach, dispersion_k

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="waves.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:void|int|int64_t|const char \*)\s*\*?\s*(wv_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = S.library()
    names = declared_symbols()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(S.EXPORTS)
    net_names = declared_symbols("waves_net.h")
    for n in net_names:
        assert hasattr(lib, n), n
    assert set(net_names) == set(S.NET_EXPORTS)


def test_net_parameter_layout_matches_oracle():
    from oracle import unet
    assert S.library().wv_net_param_count() == unet.param_count()


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        return  # a GPU is present: covered by the GPU tests
    lib = S.library()
    st = S._Settings()
    st.pc[:] = [float(v) for v in default_beach(n=(8, 8, 2)).push_constants()]
    h = ctypes.c_void_p()
    rc = lib.wv_create(ctypes.byref(st), ctypes.byref(h))
    assert rc != 0 and not h.value


def test_invalid_arguments_are_rejected():
    lib = S.library()
    h = ctypes.c_void_p()
    assert lib.wv_create(None, ctypes.byref(h)) == -1
    st = S._Settings()  # zero grid
    assert lib.wv_create(ctypes.byref(st), ctypes.byref(h)) == -1
    assert lib.wv_advance(None, 1, ctypes.c_float(0)) == -1
    assert lib.wv_request_time(None, 1.0) == -1


def test_push_constant_block_matches_reference_layout():
    s = default_beach()
    pc = s.push_constants()
    assert pc.dtype == np.float32 and pc.size == 32
    assert list(pc[0:3]) == [310, 48, 80]
    np.testing.assert_allclose(pc[4:8], [0.025, 0.025, 0.05, 1 / 120], rtol=1e-7)
    np.testing.assert_allclose(pc[8:12], [-1.75, -0.6, -2.0, 0.0])
    assert pc[13] == 6 and pc[16] == 7 and pc[17] == 1 and pc[19] == 10000
    np.testing.assert_allclose(pc[20:24], [0.08, dispersion_k(1.4, 0.45), 2 * np.pi / 1.4, 0.45], rtol=1e-6)
    assert pc[24] == np.float32(1e-4) and pc[26] == 1 and pc[27] == 6  # valid band + transport cache
    np.testing.assert_allclose(pc[28:32], [-1.75, 2.5, 12.0, 4.0])


def test_host_level_plan_matches_oracle_at_every_config_size():
    """wv_plan_levels (no device) gives the oracle's 16-int level descriptors
    (water/pressure_multigrid.gd:33-69) for the narrow flume, the beach and the
    refined beach, with the level counts of the derived-size table (5, 7, 8)."""
    from oracle.oracle import lib as olib, _fp
    from waves_amd.settings import narrow_flume
    cases = ((narrow_flume(), 5), (default_beach(), 7),
             (default_beach(n=(620, 96, 160), spacing=(0.0125, 0.0125, 0.025)), 8),
             (narrow_flume(n=(64, 48, 4)), None), (narrow_flume(n=(64, 48, 4), periodic_x=True), None))
    for s, count in cases:
        g = S.plan_levels(s)
        L = olib()
        h = L.wo_create(_fp(s.push_constants()), _fp(s.options())) if s.cells < 2_000_000 else None
        if h:
            o = []
            for l in range(L.wo_levels(h)):
                m = (ctypes.c_int * 16)()
                L.wo_level_meta(h, l, m)
                o.append(list(m))
            L.wo_destroy(h)
            assert g == o, (s.n, g, o)
        if count is not None:
            assert len(g) == count, (s.n, len(g))
    coarse = [m[8] for m in S.plan_levels(default_beach())]
    assert coarse == [1190400, 297600, 37440, 4680, 600, 100, 50]


def test_telemetry_decode_masks_recoverable_failures():
    c = np.zeros(96, np.float32)
    c[7], c[29] = 20, 0
    assert S.Solver._stats(c)["failure"] == 0 and S.Solver._stats(c)["trial_failure"] == 20
    c[29] = 2
    assert S.Solver._stats(c)["failure"] == 20
    c[7] = 4
    assert S.Solver._stats(c)["failure"] == 4
