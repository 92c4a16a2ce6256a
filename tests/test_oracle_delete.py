"""The oracle's restatement of batch_delete (covertreex/algo/batch_delete.py:
105-525) on its own trees: the kept forest is unchanged, the removed rows are
gone, every other node is back in a consistent forest that keeps the
covering invariant."""

import numpy as np
import pytest

import oracle
from oracle import fp64
from oracle import lib as ol

UNSET = np.iinfo(np.int32).min


def _kids(t):
    out = []
    for i in range(t.parents.size):
        lst, c = [], t.children[i]
        while c != -1:
            lst.append(int(c))
            c = t.next[c]
        out.append(lst)
    return out


def _check_forest(t):
    kids = _kids(t)
    for i, lst in enumerate(kids):
        for c in lst:
            assert t.parents[c] == i
    assert sum(len(lst) for lst in kids) == int(np.sum(t.parents >= 0))
    assert np.array_equal(np.sort(t.order), np.arange(t.order.size))


def _descendants(t, nodes):
    kids, seen, stack = _kids(t), set(), list(nodes)
    while stack:
        u = stack.pop()
        for c in kids[u]:
            if c not in seen:
                seen.add(c)
                stack.append(c)
    return seen


@pytest.mark.parametrize("seed,nrem", [(0, 40), (1, 1), (2, 300)])
def test_residual_delete_keeps_forest(oracle_lib, seed, nrem):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((1200, 2)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=24, seed=0, kernel_type=0)
    t = oracle.build_residual_tree(x, v, p, kernel_type=0, chunk=128, grid_scale=6.0)
    rem = rng.choice(1200, nrem, replace=False)
    d = oracle.delete_rows(t, x, rem, v=v, p=p, kernel_type=0, chunk=128, grid_scale=6.0)
    n1 = 1200 - nrem
    assert d.parents.size == n1
    _check_forest(d)
    keep = np.setdiff1d(np.arange(1200), rem)           # new dataset row r -> old keep[r]
    # kept nodes: neither removed nor below a removed node; their links stay
    removed_nodes = set(t.inv_order[rem].tolist())
    gone = removed_nodes | _descendants(t, removed_nodes)
    kept_old = [u for u in range(1200) if u not in gone]
    assert len(gone) - nrem == n1 - len(kept_old)
    for new_u, old_u in enumerate(kept_old):
        assert keep[d.order[new_u]] == t.order[old_u]
        assert d.levels[new_u] == t.levels[old_u]
        op = t.parents[old_u]
        if op < 0:
            assert d.parents[new_u] == -1
        else:
            assert kept_old[d.parents[new_u]] == op
    # covering for every attached node of the new tree
    rows = x[keep][d.order]
    metric = ol.residual_metric(rows, v[keep][d.order], p[keep][d.order], 1.0, 1.0, 0)
    nodes = np.nonzero(d.parents >= 0)[0]
    dsq = ol.dist_sq_pairs(metric, nodes, d.parents[nodes])
    assert np.all(dsq <= (4.0 ** (d.levels[nodes].astype(np.float64) + 1)).astype(np.float32))


def test_euclidean_delete_and_errors(oracle_lib):
    rng = np.random.default_rng(4)
    pts = rng.random((1500, 4)).astype(np.float32)
    t = oracle.build_euclidean_tree(pts, chunk=256)
    rem = rng.choice(1500, 100, replace=False)
    d = oracle.delete_rows(t, pts, rem, chunk=256, euclidean=True)
    _check_forest(d)
    keep = np.setdiff1d(np.arange(1500), rem)
    rows = pts[keep][d.order]
    metric = ol.euclidean_metric(rows)
    nodes = np.nonzero(d.parents >= 0)[0]
    dsq = ol.dist_sq_pairs(metric, nodes, d.parents[nodes])
    assert np.all(dsq <= (4.0 ** (d.levels[nodes].astype(np.float64) + 1)).astype(np.float32))
    with pytest.raises(ValueError):
        oracle.delete_rows(t, pts, [3, 3], euclidean=True)
    with pytest.raises(ValueError):
        oracle.delete_rows(t, pts, [1500], euclidean=True)
