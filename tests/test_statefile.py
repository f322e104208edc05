"""The state file format on the host (waves_amd/statefile.py, the layout of
include/waves.h WV_STATE_VERSION): write/read round trip, checksum known
answer, rejection of corrupted and foreign files.  The device side
(wv_save_state / wv_load_state) is tested in tests/test_gpu_state.py."""
import os

import numpy as np
import pytest

from waves_amd import statefile as SF
from waves_amd.settings import narrow_flume


def test_checksum_known_answer():
    K = 0x9E3779B97F4A7C15
    # words 1 and 2 (the second zero-padded): (1 + 1) * 1 * K + (2 + 1) * 3 * K
    assert SF.checksum(b"\x01\x00\x00\x00\x02") == (11 * K) % (1 << 64)
    assert SF.checksum(b"") == 0
    data = np.random.default_rng(1).integers(0, 2 ** 32, 1000, dtype=np.uint64).astype(np.uint32).tobytes()
    ref = sum((int(w) + 1) * (2 * i + 1) * K for i, w in enumerate(np.frombuffer(data, np.uint32))) % (1 << 64)
    assert SF.checksum(data) == ref


def test_round_trip_and_rejections(tmp_path):
    s = narrow_flume(n=(8, 6, 4))
    rng = np.random.default_rng(3)
    fields = {0: rng.normal(size=100).astype(np.float32), 4: np.arange(96, dtype=np.float32),
              15: np.arange(70, dtype=np.uint32), 20: np.arange(12, dtype=np.uint64)}
    p = str(tmp_path / "s.wvs")
    SF.write(p, s.push_constants(), s.options(), s.n, fields)
    head, got = SF.read(p)
    assert head["n"] == (8, 6, 4) and head["version"] == SF.VERSION
    assert np.array_equal(head["pc"], s.push_constants()) and np.array_equal(head["opts"], s.options())
    for k, v in fields.items():
        assert got[k].dtype == v.dtype and np.array_equal(got[k], v)
    raw = bytearray(open(p, "rb").read())
    # offsets are 256-byte aligned
    _, _, _, off, _ = SF.ENTRY.unpack_from(raw, 256)
    assert off % 256 == 0
    raw[off + 5] ^= 0x40                    # one flipped bit in field 0
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        SF.read(p)
    raw[0:8] = b"NOTSTATE"
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="not a state file"):
        SF.read(p)
