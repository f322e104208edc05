"""CPU checks of the learned warm start's host side (waves_amd/learned.py):
its masked operator equals the FP64 oracle's assembled matrix, its feature
and target construction match the oracle's, and its initial parameters use
the oracle's layout."""
import numpy as np

from oracle import pressure64, unet
from waves_amd import learned


def random_system(n, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    N = nx * ny * nz
    work = np.zeros((14, N), np.float32)
    wet = rng.random(N) > 0.25
    work[6:12] = rng.random((6, N)).astype(np.float32)
    work[12] = work[6:12].sum(axis=0)
    geo = np.zeros((N, 2), np.float32)
    geo[:, 1] = wet
    return work.ravel(), geo.ravel(), wet


def test_masked_operator_matches_oracle_matrix():
    n = (7, 5, 4)
    work, geo, wet = random_system(n, 0)
    A, b, wet2 = pressure64.assemble(work, geo, n)
    assert np.array_equal(wet, wet2)
    p = np.random.default_rng(1).normal(size=wet.size)
    mine = learned.masked_operator(work, wet, p, n)
    assert np.allclose(mine[wet], A @ p[wet], rtol=1e-12, atol=1e-12)
    assert np.all(mine[~wet] == 0)


def test_features_match_oracle():
    rng = np.random.default_rng(2)
    N = 300
    F, r, p = rng.random(N), rng.normal(size=N), rng.normal(size=N)
    wet = rng.random(N) > 0.4
    assert np.allclose(learned.features(F, r, p, wet), unet.features(F, r, p, wet), atol=1e-6)
    t = learned.target(p, p + r, wet)
    assert np.isclose(np.sqrt((t[wet] ** 2).mean()), 1.0) and np.all(t[~wet] == 0)


def test_init_params_layout():
    assert np.array_equal(learned.init_params(4), unet.init_params(4))
