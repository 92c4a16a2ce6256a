"""Cover-tree invariants checked on the device (cx_tree_check_invariants):
covering, nesting, separation and forest consistency of residual and
Euclidean trees after the build, after inserts and after a delete — the
properties tests/test_oracle_tree.py checks on the oracle's trees.  The
separation pair count is compared with the pairs the oracle's identical tree
holds (same-level nodes of one batch strictly inside the level range), so the
check provably examined every pair it should."""

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu


def _same_level_pairs(levels, starts_sizes, lo=-10, hi=10):
    total = 0
    for s0, sz in starts_sizes:
        lv = levels[s0:s0 + sz]
        lv = lv[(lv > lo) & (lv < hi)]
        _, c = np.unique(lv, return_counts=True)
        total += int((c * (c - 1) // 2).sum())
    return total


@pytest.mark.parametrize("n,d,kt,gs,chunk", [(3000, 3, 1, 1.0, 512), (2500, 2, 0, 6.0, 256)])
def test_residual_tree_invariants(hip, n, d, kt, gs, chunk):
    from covertreex_amd.native import HipTree

    x = np.random.default_rng(n).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=48, seed=0, kernel_type=kt)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), chunk_size=chunk,
                               grid_whiten_scale=gs, kernel_type=kt)
    inv = t.check_invariants()
    assert inv["nodes"] == int(np.sum(t.get_parents() >= 0))
    batches = [(b["start"], b["size"]) for b in t.build_telemetry()]
    assert inv["batches"] == len(batches)
    assert inv["pairs"] == _same_level_pairs(t.get_levels(), batches)
    ot = oracle.build_residual_tree(x, v, p, kernel_type=kt, chunk=chunk, grid_scale=gs,
                                    exp_mode=oracle.EXP_CR)
    assert inv["pairs"] == _same_level_pairs(ot.levels, batches)


def test_invariants_after_insert_and_delete(hip, monkeypatch):
    from covertreex_amd.native import HipTree

    monkeypatch.setenv("COVERTREEX_HIP_CHECK_INVARIANTS", "1")
    x = np.random.default_rng(4).standard_normal((3000, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=32, seed=0, kernel_type=1)
    t = HipTree.build_residual(v[:2000], p[:2000], x[:2000], 1.0, np.ones(1, np.float32),
                               chunk_size=256, kernel_type=1)
    assert t.check_after_insert
    t.insert_residual(v[2000:], p[2000:], x[2000:], chunk_size=300)  # checked inside
    inv = t.check_invariants()
    assert inv["batches"] == len(t.build_telemetry())
    t.delete(np.arange(0, 3000, 7))                                  # checked inside
    inv = t.check_invariants()
    assert inv["nodes"] == int(np.sum(t.get_parents() >= 0))
    assert all(inv[k] == 0 for k in ("covering", "nesting", "separation", "forest"))


def test_euclidean_tree_invariants(hip):
    from covertreex_amd.euclidean import EuclideanHipTree

    pts = np.random.default_rng(9).random((5000, 4)).astype(np.float32)
    t = EuclideanHipTree.build(pts[:4000], chunk_size=512)
    t.insert(pts[4000:], chunk_size=256)
    inv = t.check_invariants()
    assert inv["nodes"] == int(np.sum(t.get_parents() >= 0))
    batches = [(b["start"], b["size"]) for b in t.build_telemetry()]
    assert inv["pairs"] == _same_level_pairs(t.get_levels(), batches)


def test_checker_reports_a_corrupted_tree(hip):
    # the hand-traced residual tree of tests/test_known_answers.py (1-D, V = 0,
    # p = 1, RBF, chunk 4, grid scale 2): clean it has no violation; then one
    # node at a time is rewritten on the device without relinking
    from covertreex_amd.native import CoverTreeInvariantError, HipTree

    y = np.array([0.0, 0.25, 0.25, 0.6, 0.125, 0.2, -0.1, 0.2], np.float32).reshape(-1, 1)
    v, p = np.zeros((8, 1), np.float32), np.ones(8, np.float32)

    def build():
        return HipTree.build_residual(v, p, y, 1.0, np.ones(1, np.float32), chunk_size=4,
                                      grid_whiten_scale=2.0, kernel_type=0)

    t = build()
    inv = t.check_invariants()
    assert all(inv[k] == 0 for k in ("covering", "nesting", "separation", "forest"))
    # node 7 (level -10, d^2 = 0 to its parent 5) re-pointed at node 3: d^2 =
    # 1 - exp(-0.4^2/2) = 0.077 > 4^-9 breaks covering, and 5's child list
    # still holds 7 while 3's does not: the forest is inconsistent
    t.debug_set_node(7, 3, -10)
    inv = t.check_invariants(raise_on_violation=False)
    assert inv["covering"] >= 1 and inv["forest"] >= 1
    with pytest.raises(CoverTreeInvariantError):
        t.check_invariants()
    # node 5 moved to level -4 beside node 4, both of batch 2: d^2 = 0.0028 <=
    # 4^-4 and the same grid cell break separation
    t = build()
    t.debug_set_node(5, 1, -4)
    inv = t.check_invariants(raise_on_violation=False)
    assert inv["separation"] >= 1
    assert inv["covering"] == 0 and inv["forest"] == 0
