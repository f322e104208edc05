"""Reader and writer of the solver's state file (include/waves.h: WV_STATE_VERSION).

The same layout wv_save_state writes and wv_load_state reads: a 256-byte
header (magic "WAVESTAT", version, field count, grid, the 32 + 8 setting
floats), one 32-byte table entry per field (id, bytes per element, bytes,
offset, checksum), each field's bytes at a 256-byte aligned offset.
Used to hand a device state to another consumer (the CPU oracle of the tests,
an external run) on identical inputs.
"""
from __future__ import annotations

import struct
from typing import Dict, Tuple

import numpy as np

MAGIC = b"WAVESTAT"
VERSION = 1
HEADER = struct.Struct("<8sII3iI32f8f")
ENTRY = struct.Struct("<IIQQQ")
#: field id -> element dtype (the audit state is uint32, the step timing uint64)
DTYPES = {15: np.uint32, 20: np.uint64}


K = np.uint64(0x9E3779B97F4A7C15)


def checksum(data: bytes) -> int:
    """Sum over the little-endian 32-bit words w_i (the last zero-padded) of
    (w_i + 1) (2 i + 1) K modulo 2^64 (wv_save_state's field checksum)."""
    pad = (-len(data)) % 4
    w = np.frombuffer(data + b"\0" * pad, np.uint32)
    h = np.uint64(0)
    step = 1 << 22
    with np.errstate(over="ignore"):
        for s in range(0, len(w), step):
            chunk = w[s:s + step].astype(np.uint64) + np.uint64(1)
            i = np.arange(s, s + len(chunk), dtype=np.uint64)
            h = h + np.sum(chunk * ((np.uint64(2) * i + np.uint64(1)) * K), dtype=np.uint64)
    return int(h)


def read(path: str, verify: bool = True) -> Tuple[dict, Dict[int, np.ndarray]]:
    """(header dict, {field id: array}) of a state file."""
    with open(path, "rb") as f:
        raw = f.read()
    magic, version, count, nx, ny, nz, hbytes, *rest = HEADER.unpack_from(raw, 0)
    if magic != MAGIC:
        raise ValueError("not a state file")
    if version != VERSION:
        raise ValueError(f"state file version {version}")
    pc, opts = np.array(rest[:32], np.float32), np.array(rest[32:40], np.float32)
    header = {"version": version, "n": (nx, ny, nz), "pc": pc, "opts": opts}
    fields = {}
    for k in range(count):
        fid, elem, nbytes, off, stored = ENTRY.unpack_from(raw, 256 + 32 * k)
        blob = raw[off: off + nbytes]
        if len(blob) != nbytes:
            raise ValueError(f"short field {fid}")
        if verify and checksum(blob) != stored:
            raise ValueError(f"checksum of field {fid}")
        fields[fid] = np.frombuffer(blob, DTYPES.get(fid, np.float32)).copy()
    return header, fields


def write(path: str, pc: np.ndarray, opts: np.ndarray, n, fields: Dict[int, np.ndarray]) -> None:
    """Write a state file from host arrays (in the field order given)."""
    ids = list(fields)
    blobs = [np.ascontiguousarray(fields[i], DTYPES.get(i, np.float32)).tobytes() for i in ids]
    at = (256 + 32 * len(ids) + 255) // 256 * 256
    table = []
    for i, b in zip(ids, blobs):
        table.append(ENTRY.pack(i, 8 if DTYPES.get(i) == np.uint64 else 4, len(b), at, checksum(b)))
        at = (at + len(b) + 255) // 256 * 256
    head = HEADER.pack(MAGIC, VERSION, len(ids), *[int(v) for v in n], 256,
                       *[float(v) for v in np.asarray(pc, np.float32)], *[float(v) for v in np.asarray(opts, np.float32)])
    head += b"\0" * (256 - len(head))
    with open(path, "wb") as f:
        f.write(head)
        for t in table:
            f.write(t)
        pos = 256 + 32 * len(ids)
        for (i, b), t in zip(zip(ids, blobs), table):
            off = ENTRY.unpack(t)[3]
            f.write(b"\0" * (off - pos))
            f.write(b)
            pos = off + len(b)
