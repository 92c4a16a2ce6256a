"""Euclidean device path (reference Rust Euclidean insert + knn_query) vs the
oracle: tree bit for bit, reference-form search bit for bit, exact search vs
FP64 brute force."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _gmm(n, d, comps, seed, sigma=0.02):
    rng = np.random.default_rng(seed)
    centres = rng.random((comps, d))
    lab = rng.integers(0, comps, n)
    return (centres[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32), centres


@pytest.mark.parametrize("n,d,chunk", [(3000, 8, 512), (2000, 3, 256), (1000, 2, 1000)])
def test_euclidean_tree_bit_exact(hip, n, d, chunk):
    from covertreex_amd.euclidean import EuclideanHipTree

    pts, _ = _gmm(n, d, 16, seed=n)
    ot = oracle.build_euclidean_tree(pts, chunk=chunk)
    ht = EuclideanHipTree.build(pts, chunk_size=chunk)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    assert np.array_equal(ht.get_si_cache().view(np.uint32), ot.si.view(np.uint32))


@pytest.mark.parametrize("pred", [False, True])
def test_euclidean_reference_search_matches_oracle(hip, pred):
    from covertreex_amd.euclidean import EuclideanHipTree

    pts, centres = _gmm(2500, 8, 16, seed=4)
    qs, _ = _gmm(300, 8, 16, seed=5)
    ot = oracle.build_euclidean_tree(pts, chunk=512)
    ht = EuclideanHipTree.build(pts, chunk_size=512)
    oi, od = oracle.knn_euclidean(ot, qs, 8, predecessor=pred)
    hi, hd = ht.knn_points(qs, 8, predecessor_mode=pred, mode="reference")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(hi, oi)


def test_euclidean_chain_known_answer(hip):
    # tests/test_rust_knn.py:11-57
    from covertreex_amd.euclidean import EuclideanHipTree

    pts = np.zeros((100, 2), np.float32)
    pts[:, 0] = np.arange(100)
    t = EuclideanHipTree.build(pts)
    for mode in ("reference", "exact"):
        idx, dist = t.knn_points(np.array([[0.5, 0.0]], np.float32), 2, mode=mode)
        assert sorted(idx[0].tolist()) == [0, 1]
        assert np.allclose(dist[0], [0.5, 0.5])


@pytest.mark.parametrize("pred", [False, True])
def test_euclidean_exact_vs_fp64(hip, pred):
    from covertreex_amd.euclidean import EuclideanHipTree

    pts, _ = _gmm(20000, 8, 64, seed=7)
    qs = pts[:500] if pred else _gmm(500, 8, 64, seed=8)[0]
    t = EuclideanHipTree.build(pts)
    k = 32
    idx, dist = t.knn_points(qs, k, predecessor_mode=pred, mode="exact")
    d64 = np.sqrt(((qs[:, None, :].astype(np.float64) - pts[None].astype(np.float64)) ** 2)
                  .sum(-1))
    if pred:
        d64 = np.where(np.arange(20000)[None, :] < np.arange(500)[:, None], d64, np.inf)
    for r in range(qs.shape[0]):
        valid = np.isfinite(d64[r])
        kk = min(k, int(valid.sum()))
        ref = np.sort(d64[r])[:kk]
        got = dist[r][:kk].astype(np.float64)
        assert np.all(np.abs(got - ref) <= 1e-6 + 1e-6 * ref)
        assert np.all(idx[r][kk:] == -1)
        kth = ref[-1] if kk else 0
        for j in set(idx[r][:kk].tolist()) ^ set(np.argsort(d64[r])[:kk].tolist()):
            assert abs(d64[r][j] - kth) <= 1e-6  # only ties at the cut may differ


def test_factory_euclidean(hip):
    import covertreex_amd as cx

    pts, _ = _gmm(3000, 4, 8, seed=1)
    tree = cx.cover_tree(pts)
    idx, dist = tree.knn(k=5, return_distances=True)
    assert idx.shape == (3000, 5)
    assert np.all(idx[:, 0] == np.arange(3000)) and np.all(dist[:, 0] == 0)


def _fp32_fold_knn(pts, qs, k, pred=False):
    """k smallest (d^2, index) pairs of the FP32 in-order fold
    d^2 = (((q0-p0)^2 + (q1-p1)^2) + ...) by brute force; sqrt of d^2 out,
    -1 / FLT_MAX past the available candidates."""
    n = pts.shape[0]
    acc = np.zeros((qs.shape[0], n), np.float32)
    for c in range(pts.shape[1]):
        df = qs[:, None, c] - pts[None, :, c]
        acc = acc + df * df
    idx = np.full((qs.shape[0], k), -1, np.int64)
    dist = np.full((qs.shape[0], k), np.finfo(np.float32).max, np.float32)
    for r in range(qs.shape[0]):
        m = r if pred else n
        order = np.lexsort((np.arange(m), acc[r, :m]))[:k]
        idx[r, :order.size] = order
        dist[r, :order.size] = np.sqrt(acc[r, order])
    return idx, dist


def _grid_points(n, d, seed):
    # small integer grid: many equal distances, so ties decide the lists
    rng = np.random.default_rng(seed)
    return rng.integers(0, 6, (n, d)).astype(np.float32)


@pytest.mark.parametrize("n,d,k,pred,kind", [
    (20000, 8, 32, False, "gmm"),
    (5000, 3, 50, True, "gmm"),
    (37, 2, 64, False, "gmm"),        # one partial block, k above n
    (4097, 5, 8, True, "grid"),       # ties everywhere, ragged last block
    (3000, 16, 64, False, "uniform"),
    (6000, 1, 10, False, "grid"),
])
def test_euclidean_exact_bitwise_vs_fp32_bruteforce(hip, n, d, k, pred, kind):
    # the box-pruned search returns exactly the brute-force list of FP32 fold
    # distances ordered by (d^2, index), bit for bit
    from covertreex_amd.euclidean import EuclideanHipTree

    rng = np.random.default_rng(n + d)
    if kind == "gmm":
        pts = _gmm(n, d, 16, seed=n)[0]
    elif kind == "grid":
        pts = _grid_points(n, d, seed=n)
    else:
        pts = rng.random((n, d)).astype(np.float32)
    if pred:
        qs = pts[:min(n, 700)]
    else:
        extra = (pts[rng.choice(n, 50)] if kind != "grid"
                 else _grid_points(50, d, seed=n + 1))
        far = (rng.standard_normal((20, d)) * 3).astype(np.float32)  # outside every box
        qs = np.concatenate([_gmm(300, d, 16, seed=n + 2)[0], extra, far])
    print(f"exact bitwise n={n} d={d} k={k} pred={pred} kind={kind}", flush=True)
    t = EuclideanHipTree.build(pts)
    idx, dist = t.knn_points(qs, k, predecessor_mode=pred, mode="exact")
    ri, rd = _fp32_fold_knn(pts, qs, k, pred=pred)
    assert np.array_equal(dist.view(np.uint32), rd.view(np.uint32))
    assert np.array_equal(idx, ri)


def test_euclidean_exact_follows_insert_and_delete(hip):
    # the Hilbert blocks are rebuilt after each change of the point set
    from covertreex_amd.euclidean import EuclideanHipTree

    rng = np.random.default_rng(11)
    pts = _gmm(3000, 4, 8, seed=11)[0]
    more = _gmm(1500, 4, 8, seed=12)[0]
    qs = _gmm(200, 4, 8, seed=13)[0]
    t = EuclideanHipTree.build(pts, chunk_size=512)
    idx, dist = t.knn_points(qs, 12, mode="exact")
    ri, rd = _fp32_fold_knn(pts, qs, 12)
    assert np.array_equal(idx, ri) and np.array_equal(dist.view(np.uint32), rd.view(np.uint32))
    t.insert(more, chunk_size=512)
    allp = np.concatenate([pts, more])
    idx, dist = t.knn_points(qs, 12, mode="exact")
    ri, rd = _fp32_fold_knn(allp, qs, 12)
    assert np.array_equal(idx, ri) and np.array_equal(dist.view(np.uint32), rd.view(np.uint32))
    rem = rng.choice(allp.shape[0], 400, replace=False)
    t.delete(rem)
    node_pts = np.delete(allp, rem, axis=0)[t.get_order()]  # node i's coordinates
    idx, dist = t.knn_points(qs, 12, mode="exact")
    ri, rd = _fp32_fold_knn(node_pts, qs, 12)
    assert np.array_equal(idx, ri) and np.array_equal(dist.view(np.uint32), rd.view(np.uint32))


@pytest.mark.parametrize("scale", [1e-3, 1.0, 300.0])
def test_euclidean_tree_near_level_descent(hip, scale):
    # levels whose radius covers the points' bounding box run in the descent
    # kernel: many of them for a small box (down to level -9), none above 9
    from covertreex_amd.euclidean import EuclideanHipTree

    pts = (_gmm(2500, 4, 8, seed=9)[0] * np.float32(scale)).astype(np.float32)
    pts[7] = pts[3]  # a duplicate point
    ot = oracle.build_euclidean_tree(pts, chunk=256)
    ht = EuclideanHipTree.build(pts, chunk_size=256)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    tel = ht.build_telemetry()
    assert [b["scan_evals"] for b in tel] == [b["filter_evals"] for b in ot.batch_stats]
    assert [b["leaders"] for b in tel] == [b["leaders"] for b in ot.batch_stats]


@pytest.mark.parametrize("fused", ["-11", "-3"])
def test_euclidean_tree_wide_levels(hip, monkeypatch, fused):
    # three launches per level for the levels above `fused` (all of them at
    # -11), one workgroup below: the same tree as the oracle's
    from covertreex_amd.euclidean import EuclideanHipTree

    monkeypatch.setenv("COVERTREEX_HIP_FUSED_LEVEL", fused)
    pts, _ = _gmm(3000, 8, 16, seed=3000)
    ot = oracle.build_euclidean_tree(pts, chunk=512)
    ht = EuclideanHipTree.build(pts, chunk_size=512)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)


@pytest.mark.parametrize("max_work", ["0", "1000", "1000000000000000"])
def test_euclidean_tree_fused_and_wide_segments(hip, monkeypatch, max_work):
    # segments of batches switch between fused and wide levels on the work
    # measured so far: the same tree and counters whatever the threshold
    from covertreex_amd.euclidean import EuclideanHipTree

    monkeypatch.setenv("COVERTREEX_HIP_FUSED_MAX_WORK", max_work)
    pts, _ = _gmm(12000, 8, 16, seed=12)
    ot = oracle.build_euclidean_tree(pts, chunk=256)
    ht = EuclideanHipTree.build(pts, chunk_size=256)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    tel = ht.build_telemetry()
    assert [b["scan_evals"] for b in tel] == [b["filter_evals"] for b in ot.batch_stats]
    assert [b["leaders"] for b in tel] == [b["leaders"] for b in ot.batch_stats]
