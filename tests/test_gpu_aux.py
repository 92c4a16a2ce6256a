"""Device Nystrom factor, Hilbert order, Luby MIS, Morton order and subtree
bounds vs the oracle."""

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,m,kt,ls", [(4096, 3, 64, 1, 1.0), (3000, 2, 128, 0, 0.7),
                                         (2000, 5, 96, 0, (0.5, 1.0, 1.5, 2.0, 0.8))])
def test_nystrom_device_vs_fp64(hip, n, d, m, kt, ls):
    from covertreex_amd.nystrom import nystrom_device
    from covertreex_amd.residual import inducing_indices

    x = np.random.default_rng(n).standard_normal((n, d))
    idx = inducing_indices(n, m, 0)
    v, p = nystrom_device(x, idx, variance=1.3, lengthscale=ls, kernel_type=kt)
    v32, p32, kd, v64, p64, ind = fp64.nystrom(x, inducing_count=m, seed=0, variance=1.3,
                                               lengthscale=ls, kernel_type=kt)
    assert np.array_equal(ind, idx)
    # same FP64 algorithm up to summation order: V agrees to a few FP32 ulps of
    # the row scale, p to the conditioning of the jittered K_mm
    assert np.max(np.abs(v - v32)) <= 1e-5 * max(1.0, float(np.abs(v32).max()))
    big = p64 > 1e-3
    assert np.max(np.abs(p[big] - p32[big]) / p32[big]) <= 1e-4
    assert np.all(p >= np.float32(1e-9))
    # Schur complement is PSD: |V_i|^2 <= var + jitter
    assert np.all((v.astype(np.float64) ** 2).sum(1) <= 1.3 * (1 + 1e-5))


def test_nystrom_rejects_non_pd(hip):
    from covertreex_amd import _ffi
    from covertreex_amd.nystrom import nystrom_device

    x = np.zeros((10, 2))  # identical points, jitter 0: singular K_mm
    with pytest.raises(_ffi.HipError):
        nystrom_device(x, np.arange(4), jitter=0.0)


@pytest.mark.parametrize("n,d", [(5000, 2), (3000, 3), (4096, 8), (100, 1), (64, 21)])
def test_hilbert_order_device(hip, n, d):
    from covertreex_amd.order import hilbert_permutation

    pts = np.random.default_rng(d).standard_normal((n, d))
    pts = np.vstack([pts, pts[:7]])  # duplicate codes: stable ties
    assert np.array_equal(hilbert_permutation(pts), oracle.hilbert_order(pts))


@pytest.mark.parametrize("seed,n,p", [(0, 200, 0.1), (1, 500, 0.05), (2, 60, 0.05), (3, 1, 0.0)])
def test_luby_mis_device(hip, seed, n, p):
    from covertreex_amd.order import run_mis

    rng = np.random.default_rng(seed)
    a = np.triu(rng.random((n, n)) < p, 1)
    a = a | a.T
    indptr = np.concatenate([[0], np.cumsum(a.sum(1))]).astype(np.int64)
    indices = np.concatenate([np.nonzero(a[i])[0] for i in range(n)]).astype(np.int64)
    sel, it = run_mis(indptr, indices, seed=11)
    want, wit = oracle.luby_mis(indptr, indices, seed=11)
    assert np.array_equal(sel.astype(bool), want) and it == wit


def test_morton_and_subtree_bounds_device(hip):
    from covertreex_amd.native import compute_subtree_bounds, hilbert_order

    x = np.random.default_rng(0).standard_normal((10000, 3)).astype(np.float32)
    x = np.vstack([x, x[:100]])
    order, _ = oracle.morton_order(x)
    assert np.array_equal(hilbert_order(x), order)
    assert np.array_equal(hilbert_order(x[:, :2]), np.arange(x.shape[0]))
    rng = np.random.default_rng(1)
    n = 5000
    parents = np.full(n, -1, np.int64)
    for i in range(1, n):
        if rng.random() < 0.9:
            parents[i] = rng.integers(0, i)
    n2d = rng.permutation(n).astype(np.int64)
    mn, mx = compute_subtree_bounds(parents, n2d)
    omn, omx = oracle.subtree_bounds(parents, n2d)
    assert np.array_equal(mn, omn) and np.array_equal(mx, omx)
