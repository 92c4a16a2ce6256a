"""Pin the oracle's tree build (src/algo/batch/mod.rs:447-800), cover radii
(src/algo.rs:52-106) and subtree bounds (src/tree.rs:144-185) by the
invariants the construction guarantees and by brute force on small trees."""

import numpy as np
import pytest

import oracle
from oracle import fp64
from oracle import lib as ol

UNSET = np.iinfo(np.int32).min


def _data(n, d, m, kt, seed=0):
    x = np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt)
    return x, v, p


def _children_lists(t):
    out = []
    for i in range(t.parents.size):
        lst, c = [], t.children[i]
        while c != -1:
            lst.append(int(c))
            c = t.next[c]
        out.append(lst)
    return out


CASES = [
    dict(n=1500, d=3, m=32, kt=1, chunk=512, gs=1.0),
    dict(n=1200, d=2, m=24, kt=0, chunk=128, gs=6.0),
    dict(n=1000, d=5, m=16, kt=0, chunk=256, gs=1.0),
]


@pytest.mark.parametrize("case", CASES)
def test_tree_invariants(oracle_lib, case):
    x, v, p = _data(case["n"], case["d"], case["m"], case["kt"])
    t = oracle.build_residual_tree(x, v, p, kernel_type=case["kt"], chunk=case["chunk"],
                                   grid_scale=case["gs"])
    n = case["n"]
    xo, vo, po = x[t.order], v[t.order], p[t.order]
    metric = ol.residual_metric(xo, vo, po, 1.0, 1.0, case["kt"])
    # exactly one node at the top level: node 0
    assert t.levels[0] == 10 and t.parents[0] == -1
    assert np.sum(t.levels == 10) == 1
    kids = _children_lists(t)
    # child lists and parent pointers describe the same forest
    for i, lst in enumerate(kids):
        for c in lst:
            assert t.parents[c] == i
    assert sum(len(lst) for lst in kids) == int(np.sum(t.parents >= 0))
    # orphans (dropped non-leaders) have no level, parent or children
    orphan = t.levels == UNSET
    assert np.all(t.parents[orphan] == -1)
    assert all(len(kids[i]) == 0 for i in np.nonzero(orphan)[0])
    roots = np.nonzero(t.parents == -1)[0]
    assert set(roots.tolist()) == {0} | set(np.nonzero(orphan)[0].tolist())
    # covering: a node inserted at level L sat within 2^(L+1) of its parent
    # (bottom-level survivors, also at level -10, within 2^-10)
    nodes = np.nonzero(t.parents >= 0)[0]
    dsq = ol.dist_sq_pairs(metric, nodes, t.parents[nodes])
    lv = t.levels[nodes].astype(np.float64)
    assert np.all(dsq <= (4.0 ** (lv + 1)).astype(np.float32))
    # separation: leaders of one batch and level are pairwise farther than 2^L
    # and not grid neighbours
    batch = np.arange(n) // case["chunk"]
    for b in np.unique(batch):
        for L in np.unique(t.levels[(batch == b) & (t.levels > -10) & (t.levels < 10)]):
            sel = np.nonzero((batch == b) & (t.levels == L))[0]
            if sel.size < 2:
                continue
            i, j = np.triu_indices(sel.size, 1)
            d = ol.dist_sq_pairs(metric, sel[i], sel[j])
            assert np.all(d > np.float32(4.0 ** L))
            adj = oracle.grid_adjacency(xo[sel], case["gs"])
            assert not np.any(adj[i, j])


@pytest.mark.parametrize("case", CASES)
def test_si_cache_and_subtree_bounds_brute_force(oracle_lib, case):
    x, v, p = _data(case["n"], case["d"], case["m"], case["kt"], seed=1)
    t = oracle.build_residual_tree(x, v, p, kernel_type=case["kt"], chunk=case["chunk"],
                                   grid_scale=case["gs"])
    kids = _children_lists(t)
    n = t.parents.size
    xo, vo, po = x[t.order], v[t.order], p[t.order]
    metric = ol.residual_metric(xo, vo, po, 1.0, 1.0, case["kt"])
    # subtree min/max of node_to_dataset by explicit descendant walks
    for node in range(n):
        stack, members = [node], []
        while stack:
            u = stack.pop()
            members.append(u)
            stack.extend(kids[u])
        vals = t.order[members]
        assert t.subtree_min[node] == vals.min() and t.subtree_max[node] == vals.max()
    # cover radius recurrence under node 0 (other roots keep 0)
    reach = np.zeros(n, bool)
    stack = [0]
    while stack:
        u = stack.pop()
        reach[u] = True
        stack.extend(kids[u])
    for node in range(n):
        if not reach[node] or not kids[node]:
            assert t.si[node] == 0.0
            continue
        c = np.array(kids[node])
        d = np.sqrt(ol.dist_sq_pairs(metric, np.full(c.size, node), c))
        assert t.si[node] == np.max(d + t.si[c])


def test_build_is_deterministic_and_exp_modes_agree_here(oracle_lib):
    x, v, p = _data(2048, 3, 64, 1, seed=2)
    a = oracle.build_residual_tree(x, v, p, kernel_type=1, exp_mode=ol.EXP_LIBM)
    b = oracle.build_residual_tree(x, v, p, kernel_type=1, exp_mode=ol.EXP_LIBM)
    c = oracle.build_residual_tree(x, v, p, kernel_type=1, exp_mode=ol.EXP_CR)
    for f in ("order", "parents", "children", "next", "levels", "si"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
        # the 1-ulp libm/correctly-rounded exp differences flip no comparison here
        assert np.array_equal(getattr(a, f), getattr(c, f))


def test_default_grid_leaves_most_points_orphaned(oracle_lib):
    # With grid_whiten_scale 1.0 on N(0,1) coordinates the grid graph inside a
    # Morton batch is nearly complete: few leaders, most far points dropped.
    x, v, p = _data(4096, 3, 64, 1)
    t = oracle.build_residual_tree(x, v, p, kernel_type=1)
    assert t.stats["orphans_dropped"] > 0.9 * 4096
    assert int(np.sum(t.parents == -1)) == t.stats["orphans_dropped"] + 1
