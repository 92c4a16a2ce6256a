"""Parity at the full sizes of the two remaining workloads.

- Euclidean: N=1,048,576 points in D=8 from a 64-component Gaussian mixture
  (centres U[0,1]^8, sigma 0.02), 65,536 separate queries, k=32, batches of 512.
  The tree is compared with the oracle bit for bit; the reference-form search
  bit for bit on a query sample; the exact search against FP64 brute force on
  a sample and through size-independent properties on every row.
- Mixed-precision residual: N=262,144, D=5, U[-3,3]^5, anisotropic RBF with
  per-dimension lengthscales log-uniform in [0.5, 2], Nystrom rank 512, k=30,
  predecessor mode, all points queried; checked against FP64 brute force on a
  sample of rows (sets equal except ties within 1e-6, distance error <= 1e-5,
  every error bound covering the actual error) and through properties on all
  rows.
"""

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu


def _mixture(n, seed, centres, sigma=0.02):
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, centres.shape[0], n)
    return (centres[lab] + sigma * rng.standard_normal((n, centres.shape[1]))).astype(np.float32)


def test_euclidean_million_points(hip):
    from covertreex_amd.euclidean import EuclideanHipTree

    centres = np.random.default_rng(0).random((64, 8))
    pts = _mixture(1 << 20, 0, centres)
    qs = _mixture(65536, 1, centres)
    k = 32
    print("euclidean 1M: oracle build", flush=True)
    ot = oracle.build_euclidean_tree(pts, chunk=512)
    print("euclidean 1M: device build", flush=True)
    ht = EuclideanHipTree.build(pts, chunk_size=512)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_levels(), ot.levels)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)

    print("euclidean 1M: reference-form search on a sample", flush=True)
    sample = np.random.default_rng(2).choice(qs.shape[0], 1024, replace=False)
    oi, od = oracle.knn_euclidean(ot, qs[sample], k)
    ri, rd = ht.knn_points(qs[sample], k, mode="reference")
    assert np.array_equal(rd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(ri, oi)

    print("euclidean 1M: exact search, all queries", flush=True)
    idx, dist = ht.knn_points(qs, k, mode="exact")
    assert np.all(idx >= 0) and np.all(idx < pts.shape[0])
    assert np.all(np.diff(dist, axis=1) >= 0)
    # each row's distances are the FP32 distances of its own indices
    chk = np.sqrt(((qs[:, None, :].astype(np.float64) - pts[idx].astype(np.float64)) ** 2)
                  .sum(-1))
    assert np.max(np.abs(chk - dist)) <= 1e-6
    # FP64 brute force on a sample: same sets up to ties at the cut
    fs = sample[:128]
    p64 = pts.astype(np.float64)
    for r in fs:
        d64 = np.sqrt(((p64 - qs[r].astype(np.float64)) ** 2).sum(-1))
        part = np.argpartition(d64, k)[:k + 1]
        ref = part[np.argsort(d64[part])]
        kth = d64[ref[k - 1]]
        assert np.all(np.abs(dist[r] - np.sort(d64[ref[:k]])) <= 1e-6)
        for j in set(idx[r].tolist()) ^ set(ref[:k].tolist()):
            assert abs(d64[j] - kth) <= 1e-6


def test_mixed_262k_ard_rbf(hip):
    from covertreex_amd.native import HipTree
    from covertreex_amd.nystrom import nystrom_device
    from covertreex_amd.residual import inducing_indices

    n, d, m, k = 262144, 5, 512, 30
    x = np.random.default_rng(0).uniform(-3, 3, (n, d)).astype(np.float32)
    ls = np.exp(np.random.default_rng(0).uniform(np.log(0.5), np.log(2.0), d))
    print("mixed 262k: Nystrom", flush=True)
    v, p = nystrom_device(x.astype(np.float64), inducing_indices(n, m, 0), variance=1.0,
                          lengthscale=ls, kernel_type=0)
    ls32 = ls.astype(np.float32)
    print("mixed 262k: build", flush=True)
    t = HipTree.build_residual(v, p, x, 1.0, ls32, kernel_type=0)
    q = np.arange(n)
    print("mixed 262k: kNN", flush=True)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=True, mode="mixed",
                                   return_error=True)
    valid = idx >= 0
    assert np.array_equal(valid.sum(1), np.minimum(k, q))
    assert np.all(np.where(valid, idx, -1) < q[:, None])
    assert np.all(np.diff(np.where(valid, dst, 2.0), axis=1) >= 0)
    assert np.all(np.where(valid, err, 0) <= 1e-5)

    print("mixed 262k: FP64 sample", flush=True)
    rng = np.random.default_rng(1)
    sample = np.concatenate([rng.choice(n, 96, replace=False), [1, 2, 29, 30, 31, n - 1]])
    ls64 = ls32.astype(np.float64)
    v64 = v.astype(np.float64)  # converted once, not per call
    bi, bd, _ = fp64.brute_knn_fp64(x, v64, p, sample, k, predecessor=True, kernel_type=0,
                                    lengthscale=ls64, block=16)
    worst = 0.0
    for r, qi in enumerate(sample):
        got = idx[qi][idx[qi] >= 0]
        ref = bi[r][bi[r] >= 0]
        assert got.size == ref.size
        if got.size == 0:
            continue
        d64 = fp64.residual_dist_fp64(x, v64, p, [qi], got, kernel_type=0, lengthscale=ls64)[0]
        e = np.abs(dst[qi][:got.size].astype(np.float64) - d64)
        worst = max(worst, float(e.max()))
        assert np.all(e <= 1e-5)
        assert np.all(e <= err[qi][:got.size].astype(np.float64) + 2e-7)
        if set(got.tolist()) != set(ref.tolist()):
            kth = bd[r][ref.size - 1]
            diff = np.array(sorted(set(got.tolist()) ^ set(ref.tolist())))
            dd = fp64.residual_dist_fp64(x, v64, p, [qi], diff, kernel_type=0,
                                         lengthscale=ls64)[0]
            assert np.all(np.abs(dd - kth) <= 1e-6)
    print(f"mixed 262k: sample max |d - d64| = {worst:.3e}", flush=True)
