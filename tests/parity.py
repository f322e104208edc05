"""Parity harness: the HIP solver against the CPU oracle on identical inputs.

Test infrastructure.  Before every op group the oracle's state is copied into
the GPU solver, so each group is checked on bit-identical inputs; after the
group every observable field is compared.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from waves_amd import solver as S

#: fields copied oracle -> GPU before a group (the whole per-step state)
STATE_FIELDS = (O.FIELD_BANK_A, O.FIELD_BANK_B, O.FIELD_WORK, O.FIELD_GEO, O.FIELD_CONTROL,
                O.FIELD_PLANE, O.FIELD_MOTION, O.FIELD_HISTORY, O.FIELD_COASTAL_HISTORY)
COMPARE_FIELDS = STATE_FIELDS + (O.FIELD_IMG_PLANES, O.FIELD_IMG_KINDS, O.FIELD_PROBE)
NAMES = {0: "bankA", 1: "bankB", 2: "work", 3: "geo", 4: "control", 5: "plane", 6: "motion",
         7: "img_planes", 8: "img_kinds", 9: "history", 10: "coastal_history", 11: "probe",
         13: "mg_vectors", 14: "snapshot", 15: "audit", 16: "mg_rows", 17: "mg_linv"}

#: one batch of the coupled step program as op groups (water/coupled_gpu.gd:38-56,
#: water/coastal_gpu.gd:28-33).  A group is compared as a unit: the GPU fuses a
#: reduction op with its single-group finaliser.  "build", "cycle0" and
#: "cycle1" are the multigrid hooks of water/pressure_gpu.gd:70-74.
def batch_groups(iterations: int = 8):
    g = [(43,), (1,), (20, 21), (45, 46), (22,)]
    for _ in range(3):
        g += [(23,), (24,), (55, 56), (25, 26)]
    g += [(27,), (28,), (80, 81), (82,), (83, 84), (87, 88), (17,), (18,), (16,), (2,), (47,), (63,), (64,),
          ("build",), (3, "cycle0", 4)]
    for _ in range(iterations):
        g += [(5,), (6, 7), (8, "cycle1", 9), (10,)]
    g += [(11,), (65,), (47,), (30,), (31,), (32,), (33,), (34,), (47,), (48,), (35, 36), (62, 66), (12, 13),
          (44,), (19,), (85,), (89,)]
    return g


def run_oracle(o: O.Oracle, group):
    for op in group:
        if op == "build":
            o.mg_record(True)
        elif op == "cycle0":
            o.mg_record(False, True)
        elif op == "cycle1":
            o.mg_record(False, False)
        else:
            o.op(op)


def run_gpu(g: S.Solver, group):
    for op in group:
        if op == "build":
            g.mg_record(True)
        elif op in ("cycle0", "cycle1"):
            g.mg_record(False, op == "cycle0")
        else:
            g.op(op)
    g.synchronize()


def sync_state(o: O.Oracle, g: S.Solver, fields=STATE_FIELDS):
    for f in fields:
        g.write(f, o.get(f))


def sync_mg(o: O.Oracle, g: S.Solver):
    g.write(S.FIELD_MG_VECTORS, o.get(O.FIELD_MG_VECTORS))
    g.write(S.FIELD_MG_ROWS, np.ascontiguousarray(o.get(O.FIELD_MG_ROWS)[0::2]))
    g.write(S.FIELD_MG_LINV, o.get(O.FIELD_MG_LINV))
    g.write(S.FIELD_SNAPSHOT, o.get(O.FIELD_SNAPSHOT))
    g.write(S.FIELD_AUDIT, o.get(O.FIELD_AUDIT, np.uint32))


def field_diff(a: np.ndarray, b: np.ndarray):
    """(number of elements whose bits differ, max |a-b|, max |a-b| / max(|a|, 1e-30))."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)[: a.size]
    bits = (a.view(np.uint32) != b.view(np.uint32)) & ~(np.isnan(a) & np.isnan(b))
    nd = int(bits.sum())
    if nd == 0:
        return 0, 0.0, 0.0
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    d = np.where(np.isnan(d), np.inf, d)
    scale = max(float(np.nanmax(np.abs(a.astype(np.float64)))), 1e-30)
    return nd, float(d.max()), float(d.max()) / scale


def compare(o: O.Oracle, g: S.Solver, fields=COMPARE_FIELDS):
    out = {}
    for f in fields:
        out[NAMES[f]] = field_diff(o.get(f), g.read(f))
    return out


def make_pair(settings):
    o = O.Oracle(settings)
    g = S.Solver(settings)
    return o, g


def start_pair(o: O.Oracle, g: S.Solver):
    o.initial()
    g.start()


def coarse_edge_entries(o: O.Oracle):
    """Per coarse level l >= 1: (first row entry, count, width, mask[width, count])
    marking the row entries whose stencil offset leaves the level through a
    non-periodic face, where no coarse node exists.  The offsets are read off a
    central row of the oracle's (value, column) rows; every entry with a nonzero
    coefficient is checked to point at the wrapped target those offsets give."""
    rows = o.get(O.FIELD_MG_ROWS).reshape(-1, 2)
    cols, vals = rows[:, 1].view(np.int32), rows[:, 0]
    perx = bool(o.settings.periodic_x)
    out = []
    for l, meta in enumerate(o.levels()):
        if l == 0:
            continue
        nx, ny, nz = meta[:3]
        N, W, r0 = meta[8], meta[11], meta[12]
        q0 = np.array([nx // 2, ny // 2, nz // 2])
        i0 = int(q0[0] + nx * (q0[1] + ny * q0[2]))
        i = np.arange(N)
        q = np.stack([i % nx, (i // nx) % ny, i // (nx * ny)], 1)
        mask = np.zeros((W, N), bool)
        for k in range(W):
            c = int(cols[r0 + k * N + i0])
            d = np.array([c % nx, (c // nx) % ny, c // (nx * ny)]) - q0
            d[2] = (d[2] + nz // 2) % nz - nz // 2
            if perx:
                d[0] = (d[0] + nx // 2) % nx - nx // 2
            t = q + d
            t[:, 2] %= nz
            if perx:
                t[:, 0] %= nx
            out_k = (t[:, 0] < 0) | (t[:, 0] >= nx) | (t[:, 1] < 0) | (t[:, 1] >= ny)
            target = t[:, 0] + nx * (t[:, 1] + ny * t[:, 2])
            e = r0 + k * N + i
            live = ~out_k & (vals[e] != 0)
            assert np.array_equal(cols[e][live], target[live]), ("stencil offsets not recovered", l, k)
            mask[k] = out_k
        out.append((r0, N, W, mask))
    return out
