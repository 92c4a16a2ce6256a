"""Incremental insert on device-resident trees (CoverTreeWrapper.insert_residual
/ insert, src/lib.rs:207-300) and the build / query telemetry, against the
oracle: the grown tree bit for bit, kNN on it, per-batch counters equal to the
oracle's per-batch statistics, and query counters equal to the oracle's
distance-evaluation counts."""

import json

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu


def _data(n, d, m, kt, seed):
    x = np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt)
    return x, v, p


def _same_tree(ht, ot):
    assert np.array_equal(ht.get_order(), ot.order)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    assert np.array_equal(ht.get_si_cache().view(np.uint32), ot.si.view(np.uint32))


def _same_batches(ht, ot):
    hb = ht.build_telemetry()
    assert len(hb) == len(ot.batch_stats)
    for h, o in zip(hb, ot.batch_stats):
        assert (h["start"], h["size"]) == (o["start"], o["size"])
        assert h["scan_evals"] == o["filter_evals"]
        assert h["mis_evals"] == o["conflict_pairs"]
        assert h["nonleader_evals"] == o["nonleader_evals"]
        assert h["leaders"] == o["leaders"]
        assert h["orphans"] == o["orphans_dropped"]
        assert h["far_levels"] == o["far_levels"]
        assert h["total_ns"] >= h["scan_ns"] >= 0


@pytest.mark.parametrize("n0,n1,kt,gs,ichunk", [(2048, 3000, 1, 1.0, 300),
                                                (1500, 2600, 0, 8.0, 1024),
                                                (1000, 1001, 1, 1.0, 1)])
def test_residual_insert_matches_oracle(hip, n0, n1, kt, gs, ichunk):
    from covertreex_amd.native import HipTree

    x, v, p = _data(n1, 3, 64, kt, seed=n1)
    ht = HipTree.build_residual(v[:n0], p[:n0], x[:n0], 1.0, np.ones(1, np.float32),
                                chunk_size=512, grid_whiten_scale=gs, kernel_type=kt)
    ht.insert_residual(v[n0:], p[n0:], x[n0:], chunk_size=ichunk)
    assert ht.n == n1
    ot = oracle.build_residual_tree_with_insert(
        x[:n0], v[:n0], p[:n0], x[n0:], v[n0:], p[n0:], kernel_type=kt, chunk=512,
        insert_chunk=ichunk, grid_scale=gs, exp_mode=oracle.EXP_CR)
    _same_tree(ht, ot)
    _same_batches(ht, ot)
    # kNN over the grown tree: reference form against the oracle's search
    q = np.arange(0, n1, 7)
    oi, od, _ = oracle.knn_residual(ot, q, 10, predecessor=True, budget=())
    hi, hd = ht.knn_residual(q, 10, predecessor_mode=True, mode="reference")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    # mixed engine on the grown tree (its prepared data is rebuilt)
    mi, md = ht.knn_residual(q, 10, predecessor_mode=True, mode="mixed")
    bi, bd, _ = fp64.brute_knn_fp64(x, v, p, q, 10, predecessor=True, kernel_type=kt)
    for r in range(q.size):
        got, ref = mi[r][mi[r] >= 0], bi[r][bi[r] >= 0]
        assert got.size == ref.size
        assert np.all(np.abs(md[r][: got.size] - bd[r][: ref.size]) <= 1e-5)


def test_euclidean_insert_matches_oracle(hip):
    from covertreex_amd.euclidean import EuclideanHipTree

    rng = np.random.default_rng(3)
    pts = rng.random((4000, 5)).astype(np.float32)
    ht = EuclideanHipTree.build(pts[:2500], chunk_size=512)
    ht.insert(pts[2500:], chunk_size=256)
    ot = oracle.build_euclidean_tree_with_insert(pts[:2500], pts[2500:], chunk=512,
                                                 insert_chunk=256)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    _same_batches(ht, ot)
    qs = rng.random((200, 5)).astype(np.float32)
    oi, od = oracle.knn_euclidean(ot, qs, 8)
    hi, hd = ht.knn_points(qs, 8, mode="reference")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(hi, oi)


def test_query_telemetry_counts(hip):
    from covertreex_amd.native import HipTree

    x, v, p = _data(2000, 2, 32, 0, seed=11)
    ht = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), chunk_size=64,
                                grid_whiten_scale=8.0, kernel_type=0)
    ot = oracle.build_residual_tree(x, v, p, kernel_type=0, chunk=64, grid_scale=8.0,
                                    exp_mode=oracle.EXP_CR)
    q = np.arange(2000)
    assert ht.last_query_telemetry() is None
    ht.set_query_telemetry(True)
    ht.knn_residual(q, 10, predecessor_mode=True, mode="reference")
    t = ht.last_query_telemetry()
    _, _, evals = oracle.knn_residual(ot, q, 10, predecessor=True, budget=())
    assert t["engine_mode"] == "reference-f32" and t["queries"] == 2000
    assert t["distance_evals"] == int(evals.sum())
    assert len(t["frontier_levels"]) >= 1 and sum(t["block_size_counts"].values()) > 0
    ht.knn_residual(q, 10, predecessor_mode=True, mode="mixed")
    t = ht.last_query_telemetry()
    assert t["engine_mode"] == "mixed"
    assert t["distance_evals"] == int(np.minimum(q, 2000).sum())  # every pair j < i screened
    assert t["listed_candidates"] >= 10 * 1990 - 45
    ht.clear_last_query_telemetry()
    assert ht.last_query_telemetry() is None


def test_engine_writes_batch_records(hip, tmp_path):
    from covertreex_amd.engine import HipEngine, RuntimeConfig
    from covertreex_amd.telemetry import REQUIRED_KEYS, SCHEMA_ID, BenchmarkLogWriter

    x = np.random.default_rng(5).standard_normal((1500, 3)).astype(np.float32)
    path = tmp_path / "batches.jsonl"
    with BenchmarkLogWriter(str(path), runtime={"engine": "hip"}) as w:
        tree = HipEngine().build(x, runtime=RuntimeConfig(metric="residual_correlation"),
                                 log_writer=w, residual_params={"inducing_count": 64})
    lines = [json.loads(s) for s in path.read_text().splitlines()]
    batches = tree.handle.tree.build_telemetry()
    assert len(lines) == len(batches) == 3
    for i, rec in enumerate(lines):
        assert all(key in rec for key in REQUIRED_KEYS)
        assert rec["schema_id"] == SCHEMA_ID and rec["batch_index"] == i
        assert rec["selected"] == batches[i]["leaders"]
        assert rec["candidates"] == batches[i]["size"]


@pytest.mark.parametrize("seed,nrem,gs", [(0, 60, 6.0), (1, 1, 1.0), (2, 700, 6.0)])
def test_residual_delete_matches_oracle(hip, seed, nrem, gs):
    from covertreex_amd.native import HipTree

    x, v, p = _data(2500, 2, 32, 0, seed=seed)
    ht = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), chunk_size=256,
                                grid_whiten_scale=gs, kernel_type=0)
    ot = oracle.build_residual_tree(x, v, p, kernel_type=0, chunk=256, grid_scale=gs,
                                    exp_mode=oracle.EXP_CR)
    rem = np.random.default_rng(seed + 10).choice(2500, nrem, replace=False)
    ht.delete(rem)
    od = oracle.delete_rows(ot, x, rem, v=v, p=p, kernel_type=0, chunk=256, grid_scale=gs,
                            exp_mode=oracle.EXP_CR)
    assert ht.n == 2500 - nrem
    _same_tree(ht, od)
    keep = np.setdiff1d(np.arange(2500), rem)
    xk, vk, pk = x[keep], v[keep], p[keep]
    q = np.arange(0, keep.size, 5)
    oi, odist, _ = oracle.knn_residual(od, q, 12, predecessor=False, budget=())
    hi, hd = ht.knn_residual(q, 12, predecessor_mode=False, mode="reference")
    assert np.array_equal(hd.view(np.uint32), odist.view(np.uint32))
    mi, md = ht.knn_residual(q, 12, predecessor_mode=True, mode="mixed")
    bi, bd, _ = fp64.brute_knn_fp64(xk, vk, pk, q, 12, predecessor=True, kernel_type=0)
    for r in range(q.size):
        got, ref = mi[r][mi[r] >= 0], bi[r][bi[r] >= 0]
        assert got.size == ref.size
        assert np.all(np.abs(md[r][: got.size] - bd[r][: ref.size]) <= 1e-5)
    with pytest.raises(ValueError):
        ht.delete([0, 0])


def test_euclidean_delete_matches_oracle(hip):
    from covertreex_amd.euclidean import EuclideanHipTree

    rng = np.random.default_rng(8)
    pts = rng.random((3000, 4)).astype(np.float32)
    ht = EuclideanHipTree.build(pts, chunk_size=512)
    ot = oracle.build_euclidean_tree(pts, chunk=512)
    rem = rng.choice(3000, 250, replace=False)
    ht.delete(rem)
    od = oracle.delete_rows(ot, pts, rem, chunk=512, euclidean=True)
    assert np.array_equal(ht.get_order(), od.order)
    assert np.array_equal(ht.get_parents(), od.parents)
    assert np.array_equal(ht.get_children(), od.children)
    assert np.array_equal(ht.get_next_node(), od.next)
    assert np.array_equal(ht.get_levels(), od.levels)
    qs = rng.random((150, 4)).astype(np.float32)
    oi, odist = oracle.knn_euclidean(od, qs, 6)
    hi, hd = ht.knn_points(qs, 6, mode="reference")
    assert np.array_equal(hd.view(np.uint32), odist.view(np.uint32))
    assert np.array_equal(hi, oi)
