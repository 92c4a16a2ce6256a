"""Pin the oracle's conflict-graph pieces: greedy MIS (src/algo/batch/mis.rs:9-59),
Luby MIS (covertreex/algo/_mis_numba.py:64-131) and grid adjacency
(src/algo/batch/mod.rs:212-327)."""

import numpy as np
import pytest

import oracle


def _random_graph(rng, n, p):
    a = rng.random((n, n)) < p
    a = np.triu(a, 1)
    a = a | a.T
    indptr = np.concatenate([[0], np.cumsum(a.sum(1))]).astype(np.int64)
    indices = np.concatenate([np.nonzero(a[i])[0] for i in range(n)] or [[]]).astype(np.int64)
    return a, indptr, indices


@pytest.mark.parametrize("seed", range(8))
def test_greedy_mis_is_lexicographically_first(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    a, indptr, indices = _random_graph(rng, n, rng.uniform(0.05, 0.5))
    sel = oracle.greedy_mis(indptr, indices)
    # independent
    assert not np.any(a[np.ix_(sel, sel)])
    # maximal
    for v in np.nonzero(~sel)[0]:
        assert np.any(a[v] & sel)
    # lexicographically first: v selected iff no selected neighbour precedes it
    for v in range(n):
        earlier = np.nonzero(a[v, :v] & sel[:v])[0]
        assert sel[v] == (earlier.size == 0)


@pytest.mark.parametrize("seed", range(6))
def test_luby_mis_independent_maximal_deterministic(oracle_lib, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 80))
    a, indptr, indices = _random_graph(rng, n, 0.3)  # max degree > 8 -> Luby rounds
    sel, iters = oracle.luby_mis(indptr, indices, seed=7)
    sel2, iters2 = oracle.luby_mis(indptr, indices, seed=7)
    assert np.array_equal(sel, sel2) and iters == iters2
    assert not np.any(a[np.ix_(sel, sel)])
    for v in np.nonzero(~sel)[0]:
        assert np.any(a[v] & sel)
    assert iters >= 1


def test_luby_low_degree_is_greedy(oracle_lib):
    rng = np.random.default_rng(9)
    a, indptr, indices = _random_graph(rng, 60, 0.03)
    assert np.diff(indptr).max() <= 8
    sel, iters = oracle.luby_mis(indptr, indices, seed=3)
    assert iters == 1
    assert np.array_equal(sel, oracle.greedy_mis(indptr, indices))


def test_grid_adjacency_known_answers(oracle_lib):
    # D <= 3: same cell or Chebyshev-1 neighbour cells
    pts2 = np.array([[0.1, 0.1], [0.9, 0.2], [1.5, 1.5], [2.5, 0.1], [-0.5, -0.5]], np.float32)
    adj = oracle.grid_adjacency(pts2, 1.0)
    want = np.array([
        [0, 1, 1, 0, 1],
        [1, 0, 1, 0, 1],
        [1, 1, 0, 1, 0],
        [0, 0, 1, 0, 0],
        [1, 1, 0, 0, 0],
    ], bool)
    assert np.array_equal(adj, want)
    # D > 3: only identical cells (neighbour keys span 3 axes only)
    pts5 = np.array([[0.1] * 5, [0.9] * 5, [1.1] + [0.1] * 4], np.float32)
    adj5 = oracle.grid_adjacency(pts5, 1.0)
    assert np.array_equal(adj5, np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], bool))
    # scale shrinks cells
    adjs = oracle.grid_adjacency(pts2, 4.0)
    assert not adjs[0, 1] and not adjs[0, 2]
