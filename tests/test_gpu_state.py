"""GPU: the state file through the C-ABI (wv_save_state / wv_load_state):
a loaded solver continues bit-identically; the file the library writes reads
back on the host with the same fields; a file written on the host loads into
the device; foreign settings are rejected; a device state handed to the CPU
oracle through the file steps identically (the identical-input harness)."""
import numpy as np
import pytest

from waves_amd import solver as S
from waves_amd import statefile as SF
from waves_amd.settings import narrow_flume
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
FIELDS = (S.FIELD_BANK_A, S.FIELD_BANK_B, S.FIELD_WORK, S.FIELD_GEO, S.FIELD_CONTROL, S.FIELD_MOTION,
          S.FIELD_MG_VECTORS, S.FIELD_MG_ROWS, S.FIELD_SNAPSHOT)


def test_loaded_solver_continues_identically(gpu, tmp_path):
    s = narrow_flume(history_capacity=200)
    a = S.Solver(s)
    a.start()
    a.run_steps(30)
    p = str(tmp_path / "t30.wvs")
    a.save_state(p)
    a.run_steps(20)
    b = S.Solver(s)
    b.start()
    b.load_state(p)
    assert np.array_equal(b.control, SF.read(p)[1][S.FIELD_CONTROL])
    b.run_steps(20)
    for f in FIELDS + (S.FIELD_HISTORY, S.FIELD_IMG_KINDS):
        assert np.array_equal(a.read(f).view(np.uint32), b.read(f).view(np.uint32)), f
    a.close()
    b.close()


def test_file_matches_fields_and_host_written_file_loads(gpu, tmp_path):
    s = narrow_flume(n=(64, 48, 4), history_capacity=64)
    g = S.Solver(s)
    g.start()
    g.run_steps(3)
    p = str(tmp_path / "g.wvs")
    g.save_state(p)
    head, fields = SF.read(p)
    assert head["n"] == (64, 48, 4)
    for f, v in fields.items():
        dev = g.read(f, v.dtype)
        assert np.array_equal(dev, v), f
    fields[S.FIELD_BANK_A] = fields[S.FIELD_BANK_A].copy()
    fields[S.FIELD_BANK_A][:10] = 0.5
    q = str(tmp_path / "h.wvs")
    SF.write(q, head["pc"], head["opts"], head["n"], fields)
    g.load_state(q)
    assert np.array_equal(g.read(S.FIELD_BANK_A), fields[S.FIELD_BANK_A])
    other = S.Solver(narrow_flume(n=(64, 48, 4), history_capacity=64, amplitude=0.05))
    with pytest.raises(S.WavesError, match="settings differ"):
        other.load_state(q)
    raw = bytearray(open(q, "rb").read())
    raw[-3] ^= 1
    open(q, "wb").write(bytes(raw))
    with pytest.raises(S.WavesError, match="checksum"):
        g.load_state(q)
    g.close()
    other.close()


def test_device_state_steps_identically_on_the_oracle(gpu, tmp_path):
    """A device state after 25 steps, handed to the oracle through the file
    (its multigrid rebuilt from the loaded operator on both sides): one more
    accepted step on each agrees to max |dF| <= 1e-4 with identical acceptance
    and bit-exact cell flags."""
    s = narrow_flume(history_capacity=64)
    g = S.Solver(s)
    g.start()
    g.run_steps(25)
    audit = g.read(S.FIELD_AUDIT, np.uint32)
    audit[0] = 0
    g.write(S.FIELD_AUDIT, audit)
    p = str(tmp_path / "h.wvs")
    g.save_state(p)
    _, fields = SF.read(p)
    o = Oracle(s)
    for f, v in fields.items():
        if f in (S.FIELD_MG_ROWS, S.FIELD_MG_LINV, S.FIELD_STEP_TIMES):
            continue      # the oracle keeps its own coarse operator and no timing
        o.set(f, v)
    o.step()
    g.step()
    co, cg = o.control, g.control
    assert co[2] == cg[2] == 26 and co[3] == cg[3] and co[1] == cg[1]
    acc = int(co[0])
    dF = float(np.abs(o.bank(acc)[: o.N] - g.read(acc)[: o.N]).max())
    print("max|dF| after one step from the loaded state:", dF, flush=True)
    assert dF <= 1e-4
    assert np.array_equal(o.get(3)[1::2], g.read(S.FIELD_GEO)[1::2])
    o.close()
    g.close()
