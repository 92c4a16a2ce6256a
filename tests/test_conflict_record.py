"""Conflict records of batch inserts and the reference's per-batch payload
(BatchInsertTelemetry -> emit_rust_batch, src/algo/batch/mod.rs:802-956,
src/lib.rs:1274-1371): parents, levels, selected, dominated and the conflict
CSR in batch-local indices.

On the CPU the oracle's record of the hand-traced 8-point trees of
tests/test_known_answers.py is turned into the payload and compared with the
payload traced by hand from the reference's code; on the GPU the device's
record equals the oracle's word for word."""

import numpy as np
import pytest

import oracle
import oracle.lib as ol
from covertreex_amd.telemetry import payloads_from_record
from oracle import fp64

# Both hand-traced builds (1-D Euclidean at x = 0, 4, 4, 9 | 2, 3.5, -1.5,
# 3.5 with grid cells of width 4; residual V = 0 at y = 0, .25, .25, .6 |
# .125, .2, -.1, .2 with cells of width 1/2) have the same conflict
# structure:
#   batch 0: a level where node 3 is far alone (leader); a level with far =
#     [1, 2] at d^2 = 0 in one grid cell -> one distance and one grid
#     conflict; 1 leads, 2 is covered (dominated) and survives to the bottom
#     (selected).  Row 1 = [2 (distance), 2, 2 (grid, both directions)].
#   batch 1 (rows 4..7): far = [4, 6] beyond r but in adjacent cells -> a
#     grid conflict only; 4 leads, 6 is dropped.  Then far = [5, 7] at d^2 =
#     0 in one cell -> distance and grid; 5 leads, 7 is covered and survives.
EXPECTED = [
    dict(selected=[1, 2, 3], dominated=[2], conflict_indptr=[0, 0, 3, 6, 6],
         conflict_indices=[2, 2, 2, 1, 1, 1]),
    dict(selected=[0, 1, 3], dominated=[3], conflict_indptr=[0, 2, 5, 7, 10],
         conflict_indices=[2, 2, 3, 3, 3, 0, 0, 1, 1, 1]),
]
PARENTS = [-1, 0, 1, 0, 0, 1, -1, 5]


@pytest.fixture
def record_on(oracle_lib):
    ol.set_conflict_log(True)
    yield
    ol.set_conflict_log(False)


def _payloads(t):
    return payloads_from_record(t.conflict_log, t.batch_stats, t.parents, t.levels)


def _check_expected(pl, levels):
    assert len(pl) == 2
    for b, (got, want) in enumerate(zip(pl, EXPECTED)):
        assert got["batch_index"] == b and got["batch_size"] == 4
        for key, val in want.items():
            assert got[key] == val, (b, key)
        assert got["parents"] == PARENTS[4 * b:4 * b + 4]
        assert got["levels"] == levels[4 * b:4 * b + 4]
        assert got["conflict_grid_local_edges"] == [1, 2][b]
        assert got["conflict_grid_leaders_raw"] == 2
        # scopes: each row's distinct neighbours (build_scopes_from_edges,
        # src/algo/batch/mod.rs:16-200, reference defaults)
        assert got["scope_indptr"] == [[0, 0, 1, 2, 2], [0, 1, 2, 3, 4]][b]
        assert got["scope_indices"] == [[2, 1], [2, 3, 0, 1]][b]
        tm = got["timings"]
        assert tm["scope_chunk_scans"] == len(want["conflict_indices"])
        assert tm["scope_chunk_dedupe"] == len(want["conflict_indices"]) - [2, 4][b]
        assert tm["scope_chunk_pairs_before"] == len(want["conflict_indices"]) // 2
        assert tm["scope_chunk_max_members"] == 1 and tm["scope_chunk_pairs_after"] == 0


def test_euclidean_known_tree_payload(record_on):
    x = np.array([0.0, 4.0, 4.0, 9.0, 2.0, 3.5, -1.5, 3.5], np.float32).reshape(-1, 1)
    t = oracle.build_euclidean_tree(x, chunk=4, grid_scale=0.25)
    _check_expected(_payloads(t), [10, 1, -10, 3, 0, -2, np.iinfo(np.int32).min, -10])


def test_residual_known_tree_payload(record_on):
    y = np.array([0.0, 0.25, 0.25, 0.6, 0.125, 0.2, -0.1, 0.2], np.float32).reshape(-1, 1)
    t = oracle.build_residual_tree(y, np.zeros((8, 1), np.float32), np.ones(8, np.float32),
                                   kernel_type=0, chunk=4, grid_scale=2.0,
                                   exp_mode=oracle.EXP_CR)
    _check_expected(_payloads(t), [10, -3, -10, -2, -4, -5, np.iinfo(np.int32).min, -10])


def test_record_off_by_default(oracle_lib):
    x = np.array([0.0, 4.0, 4.0, 9.0], np.float32).reshape(-1, 1)
    assert oracle.build_euclidean_tree(x, chunk=4).conflict_log.size == 0


def test_payload_csr_is_symmetric(record_on):
    # every conflict row is mirrored, with the same multiplicity, in its
    # neighbour's row (the reference pushes both directions of each edge)
    x = np.random.default_rng(1).standard_normal((600, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=24, seed=0, kernel_type=1)
    t = oracle.build_residual_tree(x, v, p, kernel_type=1, chunk=128, grid_scale=1.0,
                                   exp_mode=oracle.EXP_CR)
    pl = _payloads(t)
    assert len(pl) == len(t.batch_stats)
    for b in pl:
        ip, ix = np.array(b["conflict_indptr"]), np.array(b["conflict_indices"])
        pairs = {}
        for u in range(b["batch_size"]):
            for v_ in ix[ip[u]:ip[u + 1]]:
                pairs[(u, int(v_))] = pairs.get((u, int(v_)), 0) + 1
        assert all(pairs.get((v_, u), 0) == c for (u, v_), c in pairs.items())
        # selected rows all got a level; dominated rows were far at some level
        lev = np.array(b["levels"])
        assert np.all(lev[b["selected"]] != np.iinfo(np.int32).min)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(kind="residual", n=4096, d=3, m=64, kt=1, chunk=512, gs=1.0),
    dict(kind="residual", n=3000, d=5, m=48, kt=0, chunk=256, gs=1.0, dist="uniform"),
    dict(kind="residual", n=6000, d=2, m=128, kt=0, chunk=512, gs=1.0, ls=1.5),
    dict(kind="euclidean", n=5000, d=8, chunk=512),
])
@pytest.mark.parametrize("max_work", ["0", "1000000000000000"])
def test_device_record_equals_oracle(hip, monkeypatch, case, max_work):
    from covertreex_amd.euclidean import EuclideanHipTree
    from covertreex_amd.native import HipTree, set_conflict_log

    monkeypatch.setenv("COVERTREEX_HIP_FUSED_MAX_WORK", max_work)
    rng = np.random.default_rng(7)
    print(f"case {case} max work {max_work}", flush=True)
    ol.set_conflict_log(True)
    set_conflict_log(True)
    try:
        if case["kind"] == "euclidean":
            centres = rng.random((16, case["d"]))
            x = (centres[rng.integers(0, 16, case["n"])]
                 + 0.02 * rng.standard_normal((case["n"], case["d"]))).astype(np.float32)
            ot = oracle.build_euclidean_tree(x, chunk=case["chunk"])
            ht = EuclideanHipTree.build(x, chunk_size=case["chunk"])
        else:
            ls = case.get("ls", 1.0)
            x = (rng.uniform(-3, 3, (case["n"], case["d"])) if case.get("dist") == "uniform"
                 else rng.standard_normal((case["n"], case["d"]))).astype(np.float32)
            v, p, *_ = fp64.nystrom(x, inducing_count=case["m"], seed=0,
                                    kernel_type=case["kt"], lengthscale=ls)
            ot = oracle.build_residual_tree(x, v, p, ls=ls, kernel_type=case["kt"],
                                            chunk=case["chunk"], grid_scale=case["gs"],
                                            exp_mode=oracle.EXP_CR)
            ht = HipTree.build_residual(v, p, x, 1.0, np.full(1, ls, np.float32),
                                        chunk_size=case["chunk"],
                                        grid_whiten_scale=case["gs"], kernel_type=case["kt"])
    finally:
        set_conflict_log(False)
        ol.set_conflict_log(False)
    assert np.array_equal(ht.get_parents(), ot.parents)
    rec = ht.conflict_log()
    assert rec.size > 0
    assert np.array_equal(rec, ot.conflict_log)
    hp = payloads_from_record(rec, ht.build_telemetry(), ht.get_parents(), ht.get_levels())
    op = _payloads(ot)
    keys = ("selected", "dominated", "conflict_indptr", "conflict_indices", "parents", "levels")
    assert [[b[k] for k in keys] for b in hp] == [[b[k] for k in keys] for b in op]
    # leaders counted by the device telemetry = leaders in the record
    tel = ht.build_telemetry()
    for b, pl in zip(tel, hp):
        assert b["conflict_ns"] >= 0
