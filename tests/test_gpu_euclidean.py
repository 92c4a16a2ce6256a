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
