"""Known answers that pin the oracle (and, through the C-ABI, the device) to the
reference's semantics, each traced by hand from the reference source:

- batch insert (src/algo/batch/mod.rs:521-800): a 1-D Euclidean build and a
  residual build with V = 0, p = var (d^2 = 1 - exp(-dx^2 / 2) exactly), both
  8 points in two batches of 4, covering a cover/child distance tie (the first
  found wins: strict <), a grid edge between far points more than r apart, a
  dropped far non-leader, prepend child order and bottom survivors at
  min_level;
- Luby MIS (covertreex/algo/_mis_numba.py:19-131) on graphs with degree 9;
- the residual query's budget ladder (src/algo.rs:648-683, 997-1219) on a
  hand-built star;
- the FP64 Nystrom factor (host_backend.py:201-218) against closed forms.
"""

import numpy as np
import pytest

import oracle
from oracle import fp64

UNSET = np.iinfo(np.int32).min

# ---- batch insert ------------------------------------------------------------------
#
# 1-D Euclidean, chunk 4, grid cells of width 4 (scale 0.25).
# Batch 1 = nodes 0..3 at x = 0, 4, 4, 9: node 0 is the root (level 10).
#   L = 9..4: every candidate within 2^L of the root.  L = 3 (r^2 = 64): node 3
#   (d^2 = 81) is far, alone -> leader at level 3 under the root.  L = 2
#   (r^2 = 16): nodes 1, 2 at d^2 = 16 <= 16 stay near the root (node 3 at 25).
#   L = 1 (r^2 = 4): nodes 1, 2 far; d^2(1,2) = 0 -> conflict; MIS keeps 1
#   (level 1, parent 0, prepended before 3); 2 is covered by leader 1 and stays
#   near it (d^2 = 0) to the bottom: level -10, parent 1.
# Batch 2 = nodes 4..7 at x = 2, 3.5, -1.5, 3.5.
#   Node 4 (x = 2) is at d^2 = 4 from the root and from its child 1: the root,
#   scanned first, keeps the cover (strict <) down to L = 1; at L = 0 it is far.
#   Node 6 (x = -1.5): d^2 = 2.25 from the root, far at L = 0 too.
#   Nodes 5, 7 (x = 3.5): child 1 (d^2 = 0.25) beats the root from L = 9 on,
#   and node 2 (also 0.25) does not displace 1; near until L = -1 (0.25 <=
#   0.25).
#   L = 0: far = [4, 6], d^2 = 12.25 > 1 but cells 0 and -1 are adjacent: a
#   grid edge.  MIS keeps 4 (level 0, parent 0, prepended: 0's list 4, 1, 3);
#   6 has no leader within 1 and is dropped (no parent, no level).
#   L = -2 (r^2 = 1/16): far = [5, 7], d^2 = 0 -> 5 leads (level -2, parent 1,
#   1's list 5, 2); 7 follows 5 to the bottom: level -10, parent 5.
EUCLID_X = [0.0, 4.0, 4.0, 9.0, 2.0, 3.5, -1.5, 3.5]
TREE_PARENTS = [-1, 0, 1, 0, 0, 1, -1, 5]
EUCLID_LEVELS = [10, 1, -10, 3, 0, -2, UNSET, -10]
TREE_CHILDREN = [4, 5, -1, -1, -1, 7, -1, -1]   # first child of each node
TREE_NEXT = [-1, 3, -1, -1, 1, 2, -1, -1]       # next sibling

# Residual, RBF, lengthscale 1, var 1, V = 0 (one column), p = 1, 1-D,
# chunk 4, grid scale 2 (cells of width 1/2).  d^2 = 1 - exp(-dx^2/2) < 1, so
# every candidate is near at L >= 0; the radii that split are r^2 = 4^L, L < 0.
#   Batch 1 at y = 0, 0.25, 0.25, 0.6: d^2(1,0) = 0.0308, d^2(3,0) = 0.1647,
#   d^2(1,3) = 0.0594.  L = -2 (0.0625): node 3 far, leader at -2 under 0.
#   L = -3 (0.015625): nodes 1, 2 far (from 0 and from child 3); 1 leads
#   (level -3, parent 0, prepended before 3); 2 follows 1 (d^2 = 0): level -10.
#   Batch 2 at y = 0.125, 0.2, -0.1, 0.2: node 4 is equidistant from 0 and 1
#   (d^2 = 0.00778, the same float: both coordinate gaps are 0.125): the root,
#   scanned first, keeps it; far at L = -4 (0.0039).  Node 6: 0.00499 from the
#   root, far at L = -4.  Nodes 5, 7: d^2 = 0.00125 to node 1 (beating the
#   root's 0.0198, and tied with node 2), near to L = -4, far at L = -5.
#   L = -4: far = [4, 6], d^2 = 0.025 > 0.0039, cells 0 and -1 adjacent:
#   4 leads (level -4, parent 0), 6 is dropped.  L = -5: 5 leads (parent 1),
#   7 follows 5 to the bottom.
RESID_Y = [0.0, 0.25, 0.25, 0.6, 0.125, 0.2, -0.1, 0.2]
RESID_LEVELS = [10, -3, -10, -2, -4, -5, UNSET, -10]


def _resid_inputs():
    y = np.array(RESID_Y, np.float32).reshape(-1, 1)
    return y, np.zeros((8, 1), np.float32), np.ones(8, np.float32)


def test_euclidean_batch_insert_known_tree(oracle_lib):
    x = np.array(EUCLID_X, np.float32).reshape(-1, 1)
    t = oracle.build_euclidean_tree(x, chunk=4, grid_scale=0.25)
    assert t.parents.tolist() == TREE_PARENTS
    assert t.levels.tolist() == EUCLID_LEVELS
    assert t.children.tolist() == TREE_CHILDREN
    assert t.next.tolist() == TREE_NEXT
    assert (t.min_level, t.max_level) == (-10, 10)
    assert [b["orphans_dropped"] for b in t.batch_stats] == [0, 1]
    assert [b["leaders"] for b in t.batch_stats] == [2, 2]


@pytest.mark.parametrize("exp_mode", [oracle.EXP_LIBM, oracle.EXP_CR])
def test_residual_batch_insert_known_tree(oracle_lib, exp_mode):
    y, v, p = _resid_inputs()
    t = oracle.build_residual_tree(y, v, p, kernel_type=0, chunk=4, grid_scale=2.0,
                                   exp_mode=exp_mode)
    assert t.order.tolist() == list(range(8))  # D != 3: natural order
    assert t.parents.tolist() == TREE_PARENTS
    assert t.levels.tolist() == RESID_LEVELS
    assert t.children.tolist() == TREE_CHILDREN
    assert t.next.tolist() == TREE_NEXT


@pytest.mark.gpu
def test_device_residual_batch_insert_known_tree(hip):
    from covertreex_amd.native import HipTree

    y, v, p = _resid_inputs()
    t = HipTree.build_residual(v, p, y, 1.0, np.ones(1, np.float32), chunk_size=4,
                               grid_whiten_scale=2.0, kernel_type=0)
    assert t.get_parents().tolist() == TREE_PARENTS
    assert t.get_levels().tolist() == RESID_LEVELS
    assert t.get_children().tolist() == TREE_CHILDREN
    assert t.get_next_node().tolist() == TREE_NEXT


# ---- Luby MIS ----------------------------------------------------------------------
#
# Priorities for seed 0 (x = node ^ 0 ^ it * 0x9E3779B9 mod 2^32, then x ^= x << 13,
# x ^= x >> 17, x ^= x << 5), round 1:
#   0:510c4619 1:51086638 2:5104065b 3:5100267a 4:511cc69d 5:5118e6bc
#   6:511486df 7:5110a6fe 8:512d4711 9:51296730 10:51250753 11:51212772
# Graph A: star 0 - {1..9} (degree 9 > 8: Luby rounds) plus 9 - 10 - 11.
#   Round 1 winners: 4..8 (above the centre), 9 (above 0 and 10); 0 and 10 are
#   dominated.  Round 2: 1, 2, 3, 11 have no active neighbour and win.
# Graph B: star 8 - {0..7, 9}: 8 has the largest round-1 priority of 0..9 and
#   wins alone; every leaf is dominated.


def _csr(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    indptr = np.cumsum([0] + [len(x) for x in adj]).astype(np.int64)
    indices = np.array([j for x in adj for j in sorted(x)], np.int64)
    return indptr, indices


LUBY_CASES = [
    ("star0_path", 12, [(0, i) for i in range(1, 10)] + [(9, 10), (10, 11)],
     [1, 2, 3, 4, 5, 6, 7, 8, 9, 11], 2),
    ("star8", 10, [(8, i) for i in range(10) if i != 8], [8], 1),
]


@pytest.mark.parametrize("name,n,edges,selected,iters", LUBY_CASES)
def test_luby_known_answer(oracle_lib, name, n, edges, selected, iters):
    indptr, indices = _csr(n, edges)
    sel, it = oracle.luby_mis(indptr, indices, seed=0)
    assert np.nonzero(sel)[0].tolist() == selected
    assert it == iters


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,edges,selected,iters", LUBY_CASES)
def test_device_luby_known_answer(hip, name, n, edges, selected, iters):
    from covertreex_amd.order import run_mis

    indptr, indices = _csr(n, edges)
    sel, it = run_mis(indptr, indices, seed=0)
    assert np.nonzero(sel)[0].tolist() == selected
    assert it == iters


# ---- budget ladder -----------------------------------------------------------------
#
# A star: node 0 (the query, root, cover radius 2) with children 1..200 in list
# order; dx = 0.5 - 0.002 i for i <= 192 (later children nearer), dx = 0.001 (i - 192)
# for 193..200 (the eight nearest, at the list's tail).  V = 0, p = 1, RBF: the
# distance grows with dx.  k = 5, not predecessor mode.  The root is seeded
# (one survivor); its children are gathered with parent bound 0 - 2 < kth, so
# none is ever pruned and every evaluated child is pushed.
# Default ladder [32, 64, 96], up 0.015, down 0.002, adaptive tiles:
#   tile 1 = children 1..64 (201 active + 1 frontier node > 128: tile 64): in
#   distance order 64, 63, ..., pushed until the 32nd survivor (31 pushed), yield
#   31/64 >= 0.015: budget 64.  Tile 2 = 65..128: 32 pushed, budget 96.
#   Tile 3 = 129..160 (72 active + 1 <= 128: tile 32): all 32 pushed, reaching
#   96 survivors, and the search stops.  161 distance evaluations; the five
#   nearest pushed: the query itself and 160, 159, 158, 157.
# With up 2.0, down 1.01 and one budget of 1000 every full tile is "low
# yield": after tiles 1..64 and 65..128 the streak of two ends the level;
# 129 evaluations; nearest pushed: 0, 128, 127, 126, 125.
# (With the default thresholds and tiles of at most 64, an evaluated tile
# always pushes at least one child -- a tile is evaluated only when some child's
# parent bound is <= kth, and such a child is never pruned -- so its yield is
# >= 1/64 > 0.015: the low-yield branch needs other thresholds, which the
# reference takes from COVERTREEX_RESIDUAL_BUDGET_UP / _DOWN.)


def _budget_star():
    n = 201
    dx = np.zeros(n)
    i = np.arange(1, 193)
    dx[1:193] = 0.5 - 0.002 * i
    dx[193:] = 0.001 * (np.arange(193, 201) - 192)
    parents = np.r_[-1, np.zeros(200, np.int64)]
    children = np.full(n, -1)
    children[0] = 1
    nxt = np.full(n, -1)
    nxt[1:200] = np.arange(2, 201)
    levels = np.r_[10, np.full(200, 9)]
    si = np.zeros(n, np.float32)
    si[0] = 2.0
    return oracle.tree_from_arrays(dx.reshape(-1, 1).astype(np.float32),
                                   np.zeros((n, 1), np.float32), np.ones(n, np.float32),
                                   parents=parents, children=children, next_node=nxt,
                                   levels=levels, si=si, kernel_type=0)


@pytest.mark.parametrize("kw,want,evals", [
    (dict(budget="default"), [0, 160, 159, 158, 157], 161),
    (dict(budget=(1000,), budget_up=2.0, budget_down=1.01), [0, 128, 127, 126, 125], 129),
    (dict(budget=()), [0, 193, 194, 195, 196], 201),
])
def test_budget_ladder_known_answer(oracle_lib, kw, want, evals):
    t = _budget_star()
    idx, dist, ev = oracle.knn_residual(t, [0], 5, **kw)
    assert idx[0].tolist() == want
    assert int(ev[0]) == evals
    dx = np.array([0.0] + [0.5 - 0.002 * j for j in range(1, 193)] +
                  [0.001 * j for j in range(1, 9)])
    # d^2 = 1 - exp(-dx^2/2) in the reference's FP32 arithmetic: the cancellation
    # leaves a few ulps of 1 in d^2 (large relative errors at d ~ 1e-3)
    d2_64 = 1.0 - np.exp(-0.5 * dx[want].astype(np.float32).astype(np.float64) ** 2)
    assert np.allclose(dist[0].astype(np.float64) ** 2, d2_64, rtol=0, atol=3e-7)


# ---- FP64 Nystrom ------------------------------------------------------------------
#
# With A = K_mm + eps I (eps = 1e-6 var) = L L^T and V = K_xm L^-T:
#   V V^T = K_xm A^-1 K_mx (any solve of A);
#   on the inducing rows K_xm = A - eps I, so V = L - eps L^-T there and
#   p = var - |V_i|^2 = eps - eps^2 (A^-1)_ii, in (0, eps] for every inducing row.


@pytest.mark.parametrize("kt,ls,var", [(1, 1.0, 1.0), (0, 0.7, 2.5),
                                       (0, np.array([0.6, 1.3, 0.9]), 1.0)])
def test_nystrom_closed_forms(kt, ls, var):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((400, 3))
    v32, p32, _, v, p, ind = fp64.nystrom(x, inducing_count=40, seed=1, kernel_type=kt,
                                          variance=var, lengthscale=ls)
    eps = 1e-6 * var
    kxm = fp64.kernel_fp64(x, x[ind], variance=var, lengthscale=ls, kernel_type=kt)
    a = kxm[ind] + eps * np.eye(ind.size)
    w, q = np.linalg.eigh(a)                    # an independent solve of A
    ainv = (q / w) @ q.T
    g = kxm @ ainv @ kxm.T
    assert np.max(np.abs(v @ v.T - g)) <= 1e-9 * var
    lmat = np.linalg.cholesky(a)
    assert np.max(np.abs(v[ind] - (lmat - eps * np.linalg.inv(lmat).T))) <= 1e-9 * var
    p_ind = eps - eps ** 2 * np.diag(ainv)
    assert np.all(p_ind > 0) and np.all(p_ind <= eps)
    assert np.allclose(p[ind], np.maximum(p_ind, 1e-9), rtol=1e-6, atol=1e-12)
    # every row: p = max(var - |V_i|^2, 1e-9), and the Schur complement is PSD
    assert np.all(p >= 1e-9)
    assert np.allclose(p, np.maximum(var - np.sum(v * v, axis=1), 1e-9), rtol=0, atol=1e-12)
    assert np.all(var - np.sum(v * v, axis=1) >= -1e-9 * var)
    # stored as float32 from the FP64 values
    assert np.array_equal(v32, v.astype(np.float32)) and np.array_equal(p32, p.astype(np.float32))
