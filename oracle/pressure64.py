"""FP64 pressure oracle (test infrastructure): the masked 7-point system of
pressure_core.glslinc:120-125 assembled in double from the face coefficients
and right-hand side the solver formed (work fields 6-12 and 5), solved by
scipy in FP64 (sparse direct for small systems, CG to 1e-14 from a given start
for large ones)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def assemble(work: np.ndarray, geo: np.ndarray, n):
    nx, ny, nz = (int(v) for v in n)
    N = nx * ny * nz
    work = work.reshape(14, N).astype(np.float64)
    wet = geo.reshape(N, 2)[:, 1] > 0.5
    idx = -np.ones(N, np.int64)
    idx[wet] = np.arange(wet.sum())
    cells = np.flatnonzero(wet)
    x, y, z = cells % nx, (cells // nx) % ny, cells // (nx * ny)
    # diagonal: the sum of the six face coefficients in FP64 (Dirichlet faces
    # to dry cells included), the operator the device's FP64 residual applies
    rows, cols, vals = [idx[cells]], [idx[cells]], [work[6:12, cells].sum(axis=0)]
    face = 0
    for axis in range(3):
        for sign in (-1, 1):
            q = [x.copy(), y.copy(), z.copy()]
            q[axis] = q[axis] + sign
            q[2] %= nz
            ok = (q[0] >= 0) & (q[0] < nx) & (q[1] >= 0) & (q[1] < ny)
            j = np.where(ok, q[0] + nx * (q[1] + ny * q[2]), 0)
            ok &= wet[j]
            rows.append(idx[cells[ok]]); cols.append(idx[j[ok]]); vals.append(-work[6 + face, cells[ok]])
            face += 1
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(cells), len(cells)))
    b = work[5, cells]
    return A, b, wet


def solve(A, b, x0=None, rtol=1e-14, maxiter=20000):
    if A.shape[0] <= 60000:
        return spla.spsolve(A.tocsc(), b)
    d = A.diagonal()
    M = sp.diags(1.0 / d)
    x, info = spla.cg(A, b, x0=x0, rtol=rtol, atol=0.0, maxiter=maxiter, M=M)
    return x
