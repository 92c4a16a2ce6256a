"""Pin the oracle's kNN (single_residual_knn_query, src/algo.rs:612-1242;
single_knn_query, src/algo.rs:208-323) with the reference's own test
invariants (tests/test_predecessor_mode.py, tests/test_rust_knn.py) and brute
force."""

import numpy as np
import pytest

import oracle
from oracle import fp64
from oracle import lib as ol


def _tree(n=1200, d=3, m=32, kt=1, gs=1.0, chunk=512, seed=0):
    x = np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt)
    t = oracle.build_residual_tree(x, v, p, kernel_type=kt, grid_scale=gs, chunk=chunk)
    return x, v, p, t


@pytest.mark.parametrize("budget", [(), "default"])
def test_predecessor_invariants(oracle_lib, budget):
    x, v, p, t = _tree(n=1000, gs=1.0)
    k = 30
    idx, dist, _ = oracle.knn_residual(t, np.arange(1000), k, predecessor=True, budget=budget)
    for i in range(1000):  # test_predecessor_mode.py:26-61
        valid = idx[i][idx[i] >= 0]
        assert np.all(valid < i)
    assert np.all(idx[0] == -1) and np.all(dist[0] == np.finfo(np.float32).max)
    assert set(idx[1][idx[1] >= 0].tolist()) <= {0}
    if budget == ():  # exact search fills min(k, i) (test_predecessor_mode.py:260-324)
        cnt = (idx >= 0).sum(1)
        assert np.array_equal(cnt, np.minimum(k, np.arange(1000)))
    # ascending distances, padding at the end
    d = np.where(idx >= 0, dist, 2.0)
    assert np.all(np.diff(d, axis=1) >= 0)


@pytest.mark.parametrize("gs,pred", [(1.0, False), (1.0, True), (8.0, True), (8.0, False)])
def test_exact_mode_equals_bruteforce_of_its_distances(oracle_lib, gs, pred):
    # On both the degenerate (gs=1) and the proper (gs=8) tree, the exact
    # search returns the k smallest of the f32 distances it evaluates: roots
    # through sqrt(distance_sq_idx), tree nodes through the batch variant
    # (predecessor-failed nodes never enter the result).  Compare with a brute
    # force over those per-node distances.
    x, v, p, t = _tree(n=900, gs=gs, m=16)
    k = 12
    q = np.arange(900)
    idx, dist, _ = oracle.knn_residual(t, q, k, predecessor=pred, budget=())
    xo, vo, po = x[t.order], v[t.order], p[t.order]
    metric = ol.residual_metric(xo, vo, po, 1.0, 1.0, 1)
    nodes = np.arange(900)
    is_root = t.parents == -1
    bad = 0
    for i in q:
        qn = np.full(900, t.inv_order[i])
        d = np.where(is_root, np.sqrt(ol.dist_sq_pairs(metric, qn, nodes)),
                     ol.dist_batch_pairs(metric, qn, nodes)).astype(np.float32)
        if pred:
            d[t.order >= i] = np.inf
        kk = min(k, int(np.isfinite(d).sum()))
        ref = np.sort(d)[:kk]
        if not np.array_equal(dist[i][:kk], ref):
            bad += 1
    if gs == 1.0:
        assert bad == 0  # every node is evaluated: an exhaustive search
    else:
        assert bad <= 9  # cover-radius pruning of the f32 metric may drop a few


def test_exact_mode_close_to_fp64_bruteforce(oracle_lib):
    x, v, p, t = _tree(n=1500, gs=8.0, m=24)
    k = 10
    q = np.arange(0, 1500, 7)
    idx, dist, _ = oracle.knn_residual(t, q, k, predecessor=True, budget=())
    bi, bd, bn = fp64.brute_knn_fp64(x, v, p, q, k, predecessor=True, kernel_type=1)
    agree = sum(set(idx[r][idx[r] >= 0]) == set(bi[r][bi[r] >= 0]) for r in range(q.size))
    assert agree >= 0.9 * q.size


def test_budgeted_search_is_subset_quality(oracle_lib):
    x, v, p, t = _tree(n=1500, gs=8.0, m=24)
    q = np.arange(100, 1500, 11)
    ie, de, ee = oracle.knn_residual(t, q, 16, predecessor=False, budget=())
    ib, db, eb = oracle.knn_residual(t, q, 16, predecessor=False, budget="default")
    # the budget ladder evaluates no more and finds neighbours no closer
    assert np.all(eb <= ee)
    assert np.all(db[:, 0] >= de[:, 0] - 1e-7)


def test_euclidean_chain_known_answer(oracle_lib):
    # tests/test_rust_knn.py:11-57: points on the x axis, query (0.5, 0), k=2
    pts = np.zeros((100, 2), np.float32)
    pts[:, 0] = np.arange(100)
    t = oracle.build_euclidean_tree(pts, chunk=512)
    idx, dist = oracle.knn_euclidean(t, np.array([[0.5, 0.0]], np.float32), 2)
    assert sorted(idx[0].tolist()) == [0, 1]
    assert np.allclose(dist[0], [0.5, 0.5])


def test_euclidean_equals_bruteforce(oracle_lib):
    rng = np.random.default_rng(4)
    centres = rng.random((8, 4))
    pts = (centres[rng.integers(0, 8, 3000)] + 0.05 * rng.standard_normal((3000, 4)))
    pts = pts.astype(np.float32)
    qs = (centres[rng.integers(0, 8, 200)] + 0.05 * rng.standard_normal((200, 4)))
    qs = qs.astype(np.float32)
    t = oracle.build_euclidean_tree(pts, chunk=512)
    assert np.all(t.parents[1:] >= 0)  # no grid edges, so nothing is dropped
    idx, dist = oracle.knn_euclidean(t, qs, 8)
    d = np.sqrt(((qs[:, None, :].astype(np.float64) - pts[None].astype(np.float64)) ** 2).sum(-1))
    ref = np.sort(d, axis=1)[:, :8]
    exact = np.all(np.abs(dist - ref) <= 1e-5, axis=1)
    # The batch-insert tree is not level-nested (a child may sit at a level >=
    # its parent's), so the reference's 2^(level+1) subtree bound can prune a
    # true neighbour: most rows, not all, are exact.
    lv, par = t.levels, t.parents
    assert np.any(lv[par >= 0] >= lv[par[par >= 0]])
    assert exact.sum() >= 0.9 * qs.shape[0]
    assert np.all(dist[:, 0] >= ref[:, 0] - 1e-6)
