"""ctypes bindings for the C oracle (TEST INFRASTRUCTURE, see ``oracle/__init__``)."""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

EXP_LIBM = 0
EXP_CR = 1

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_f32 = C.c_float


def build_library(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the oracle Makefile (CPU only)."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True, capture_output=True)
    return _LIB_PATH


class _QueryOpts(C.Structure):
    _fields_ = [
        ("n_budget", _i32),
        ("budget", _i64 * 8),
        ("budget_up", C.c_double),
        ("budget_down", C.c_double),
        ("stream_tile", _i32),
        ("dynamic_block", _i32),
        ("radius_floor", _f32),
        ("cap_default", _f32),
        ("threads", _i32),
    ]


class _Tree(C.Structure):
    _fields_ = [
        ("n", _i64),
        ("cap", _i64),
        ("parents", C.POINTER(_i64)),
        ("children", C.POINTER(_i64)),
        ("next", C.POINTER(_i64)),
        ("levels", C.POINTER(_i32)),
        ("si", C.POINTER(_f32)),
        ("min_level", _i32),
        ("max_level", _i32),
    ]


def build_native(out_path: str) -> bool:
    """Compile a copy tuned for the host it runs on (-O3 -march=native, same
    floating-point flags, so identical results) at ``out_path``; the CPU
    baseline times that copy.  Returns False when the compiler is missing or
    fails."""
    flags = ("-O3 -march=native -fPIC -std=c11 -fopenmp -ffp-contract=off -fno-fast-math "
             "-Wall -Wextra -Wno-unused-parameter")
    r = subprocess.run(["make", "-s", "-B", "-C", _HERE, f"OUT={out_path}", f"CFLAGS={flags}"],
                       capture_output=True)
    return r.returncode == 0 and os.path.exists(out_path)


def use_library(path: str) -> None:
    """Load the oracle from ``path`` instead of the in-tree build (before any
    other use in this process)."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if _LIB_PATH == os.path.join(_HERE, "liboracle.so"):
            build_library()
        L = C.CDLL(_LIB_PATH)
        L.ox_metric_residual.restype = _p
        L.ox_metric_residual.argtypes = [_i64, _i32, _i32, _p, _p, _p, _f32, _p, _i32, _i32, _i32]
        L.ox_metric_euclidean.restype = _p
        L.ox_metric_euclidean.argtypes = [_i64, _i32, _p]
        L.ox_metric_free.argtypes = [_p]
        L.ox_expf.restype = _f32
        L.ox_expf.argtypes = [_i32, _f32]
        L.ox_expf_range.argtypes = [_i32, C.c_uint32, _i64, _p]
        L.ox_expf_array.argtypes = [_i32, _p, _i64, _p]
        L.ox_dist_sq_pairs.argtypes = [_p, _p, _p, _i64, _p]
        L.ox_dist_batch_pairs.argtypes = [_p, _p, _p, _i64, _p]
        L.ox_morton_order.argtypes = [_i64, _i32, _p, _p, _p]
        L.ox_hilbert_codes.argtypes = [_i64, _i32, _p, _p]
        L.ox_tree_new.restype = C.POINTER(_Tree)
        L.ox_tree_new.argtypes = [_i64]
        L.ox_tree_free.argtypes = [C.POINTER(_Tree)]
        L.ox_build.argtypes = [C.POINTER(_Tree), _p, _i64, _p, _i32, _f32, _i32, _p]
        L.ox_batch_insert.argtypes = [C.POINTER(_Tree), _p, _i64, _i64, _p, _i32, _f32, _i32, _p]
        L.ox_batch_insert.restype = _i64
        L.ox_set_conflict_log.argtypes = [_p, _i64, _p, _i32]
        L.ox_si_cache.argtypes = [C.POINTER(_Tree), _p, _p]
        L.ox_subtree_bounds.argtypes = [_i64, _p, _p, _p, _p]
        L.ox_greedy_mis.argtypes = [_i64, _p, _p, _p]
        L.ox_luby_mis.restype = _i64
        L.ox_luby_mis.argtypes = [_i64, _p, _p, C.c_uint32, _p]
        L.ox_grid_edges_dense.argtypes = [_i64, _i32, _p, _f32, _p]
        L.ox_query_opts_default.argtypes = [C.POINTER(_QueryOpts)]
        L.ox_knn_residual.argtypes = [C.POINTER(_Tree), _p, _p, _p, _p, _i64, _i32, _i32, _p,
                                      C.POINTER(_QueryOpts), _p, _p, _p]
        L.ox_knn_euclidean.argtypes = [C.POINTER(_Tree), _p, _p, _i64, _i32, _i32, _p, _p, _i32]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class _Metric:
    """Owns an ``ox_metric`` over tree-ordered rows."""

    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ox_metric_free(self.h)
            self.h = None


def residual_metric(coords, v, p, var, ls, kernel_type, exp_mode=EXP_LIBM) -> _Metric:
    coords = _c(coords, np.float32)
    v = _c(v, np.float32)
    p = _c(p, np.float32)
    ls = _c(np.atleast_1d(ls), np.float32)
    n, d = coords.shape
    m = v.shape[1]
    h = lib().ox_metric_residual(n, m, d, _ptr(coords), _ptr(v), _ptr(p), float(var), _ptr(ls),
                                 ls.size, int(kernel_type), int(exp_mode))
    return _Metric(h)


def euclidean_metric(points) -> _Metric:
    pts = _c(points, np.float32)
    return _Metric(lib().ox_metric_euclidean(pts.shape[0], pts.shape[1], _ptr(pts)))


def expf(x: float, mode: int = EXP_LIBM) -> float:
    return float(lib().ox_expf(mode, float(x)))


def expf_array(x, mode: int = EXP_CR) -> np.ndarray:
    """expf of every element of a float32 array."""
    x = _c(x, np.float32)
    out = np.empty(x.shape, np.float32)
    lib().ox_expf_array(mode, _ptr(x), x.size, _ptr(out))
    return out


def expf_range(first_bits: int, count: int, mode: int = EXP_CR) -> np.ndarray:
    """expf of the `count` floats whose bit patterns follow `first_bits`."""
    out = np.empty(count, np.float32)
    lib().ox_expf_range(mode, first_bits, count, _ptr(out))
    return out


def dist_sq_pairs(metric: _Metric, a, b) -> np.ndarray:
    a = _c(a, np.int64)
    b = _c(b, np.int64)
    out = np.empty(a.size, np.float32)
    lib().ox_dist_sq_pairs(metric.h, _ptr(a), _ptr(b), a.size, _ptr(out))
    return out


def dist_batch_pairs(metric: _Metric, a, b) -> np.ndarray:
    a = _c(a, np.int64)
    b = _c(b, np.int64)
    out = np.empty(a.size, np.float32)
    lib().ox_dist_batch_pairs(metric.h, _ptr(a), _ptr(b), a.size, _ptr(out))
    return out


def morton_order(coords) -> tuple[np.ndarray, np.ndarray]:
    """hilbert_like_order (src/pcct.rs:27-67): permutation and sorted codes."""
    coords = _c(coords, np.float32)
    n = coords.shape[0]
    d = coords.shape[1] if coords.ndim == 2 else 1
    order = np.empty(n, np.int64)
    codes = np.empty(n, np.uint64)
    lib().ox_morton_order(n, d, _ptr(coords), _ptr(order), _ptr(codes))
    return order, codes


def hilbert_codes(points) -> np.ndarray:
    pts = _c(points, np.float64)
    out = np.empty(pts.shape[0], np.uint64)
    lib().ox_hilbert_codes(pts.shape[0], pts.shape[1], _ptr(pts), _ptr(out))
    return out


def hilbert_order(points) -> np.ndarray:
    """_hilbert_permutation (covertreex/algo/order/strategy.py:38-60)."""
    codes = hilbert_codes(points)
    n = codes.shape[0]
    return np.lexsort((np.arange(n, dtype=np.int64), codes)).astype(np.int64)


def greedy_mis(indptr, indices) -> np.ndarray:
    indptr = _c(indptr, np.int64)
    indices = _c(indices, np.int64)
    n = indptr.size - 1
    out = np.zeros(max(n, 1), np.uint8)
    lib().ox_greedy_mis(n, _ptr(indptr), _ptr(indices), _ptr(out))
    return out[:n].astype(bool)


def luby_mis(indptr, indices, seed: int) -> tuple[np.ndarray, int]:
    indptr = _c(indptr, np.int64)
    indices = _c(indices, np.int64)
    n = indptr.size - 1
    out = np.zeros(max(n, 1), np.uint8)
    it = lib().ox_luby_mis(n, _ptr(indptr), _ptr(indices), int(seed) & 0xFFFFFFFF, _ptr(out))
    return out[:n].astype(bool), int(it)


def grid_adjacency(coords, scale: float = 1.0) -> np.ndarray:
    coords = _c(coords, np.float32)
    n, d = coords.shape
    adj = np.zeros((n, n), np.uint8)
    lib().ox_grid_edges_dense(n, d, _ptr(coords), float(scale), _ptr(adj))
    return adj.astype(bool)


def subtree_bounds(parents, n2d) -> tuple[np.ndarray, np.ndarray]:
    parents = _c(parents, np.int64)
    n2d = _c(n2d, np.int64)
    n = parents.size
    mn = np.empty(n, np.int64)
    mx = np.empty(n, np.int64)
    lib().ox_subtree_bounds(n, _ptr(parents), _ptr(n2d), _ptr(mn), _ptr(mx))
    return mn, mx


@dataclass
class OracleTree:
    """A tree built by the oracle on tree-ordered (Morton) rows."""

    order: np.ndarray            # node -> dataset (node_to_dataset)
    inv_order: np.ndarray        # dataset -> node
    parents: np.ndarray
    children: np.ndarray
    next: np.ndarray
    levels: np.ndarray
    si: np.ndarray
    min_level: int
    max_level: int
    subtree_min: np.ndarray
    subtree_max: np.ndarray
    stats: dict = field(default_factory=dict)
    _tree: object = None
    _metric: _Metric | None = None

    def __del__(self):
        if self._tree is not None and _lib is not None:
            _lib.ox_tree_free(self._tree)
            self._tree = None


def _tree_arrays(t, n):
    tt = t.contents
    cp = lambda ptr, dt: np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)  # noqa: E731
    return (cp(tt.parents, np.int64), cp(tt.children, np.int64), cp(tt.next, np.int64),
            cp(tt.levels, np.int32), cp(tt.si, np.float32), int(tt.min_level), int(tt.max_level))


_STAT_KEYS = ["filter_evals", "conflict_pairs", "nonleader_evals", "leaders", "orphans_dropped",
              "far_levels"]


_CONFLICT_LOG = False


def set_conflict_log(enable: bool) -> None:
    """Record the conflict graph of later builds and inserts
    (``OracleTree.conflict_log``, the device library's layout)."""
    global _CONFLICT_LOG
    _CONFLICT_LOG = bool(enable)


def _insert_segments(t, metric: _Metric, segments, grid_coords, grid_scale, use_grid,
                     batch0: int = 0):
    """Chunked batch inserts over tree rows: ``segments`` is a list of
    (start, end, chunk); batches run in order and each records its stats
    (and, when turned on, its conflict record; ``batch0`` numbers the first)."""
    L = lib()
    gc = None if grid_coords is None else _c(grid_coords, np.float32)
    gd = 0 if gc is None else gc.shape[1]
    batches, logs = [], []
    total = np.zeros(8, np.int64)
    for start, end, chunk in segments:
        chunk = int(chunk) if chunk and chunk > 0 else max(end - start, 1)
        for s in range(start, end, chunk):
            e = min(s + chunk, end)
            st = np.zeros(8, np.int64)
            if _CONFLICT_LOG:
                nb, w = e - s, (e - s + 31) // 32
                cap = 21 * (4 + nb + 2 * w + 2 * nb * w) + 3 + nb
                buf = np.zeros(cap, np.int32)
                used = np.zeros(1, np.int64)
                L.ox_set_conflict_log(_ptr(buf), cap, _ptr(used), batch0 + len(batches))
            L.ox_batch_insert(t, metric.h, s, e, _ptr(gc), gd, float(grid_scale), int(use_grid),
                              _ptr(st))
            if _CONFLICT_LOG:
                L.ox_set_conflict_log(None, 0, None, 0)
                assert used[0] <= cap
                logs.append(buf[: used[0]].copy())
            total += st
            b = {k: int(v) for k, v in zip(_STAT_KEYS, st)}
            b.update(start=s, size=e - s)
            batches.append(b)
    log = np.concatenate(logs) if logs else np.zeros(0, np.int32)
    return batches, total, log


def _finish(t, metric: _Metric, n: int, order, batches, total, log=None) -> "OracleTree":
    L = lib()
    si = np.zeros(max(n, 1), np.float32)
    L.ox_si_cache(t, metric.h, _ptr(si))
    parents, children, nxt, levels, _, mn, mx = _tree_arrays(t, n)
    # the reference stores the si_cache back into the tree (src/lib.rs:1548-1561)
    C.memmove(t.contents.si, si.ctypes.data, 4 * n)
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n, dtype=np.int64)
    smin, smax = subtree_bounds(parents, order)
    tree = OracleTree(order=order, inv_order=inv, parents=parents, children=children, next=nxt,
                      levels=levels, si=si[:n].copy(), min_level=mn, max_level=mx,
                      subtree_min=smin, subtree_max=smax,
                      stats={k: int(v) for k, v in zip(_STAT_KEYS, total)}, _tree=t,
                      _metric=metric)
    tree.batch_stats = batches
    tree.conflict_log = log if log is not None else np.zeros(0, np.int32)
    return tree


def _build(metric: _Metric, n: int, order, segments, grid_coords, grid_scale, use_grid):
    """The build's chunk loop (src/lib.rs:1502-1539) and later inserts."""
    t = lib().ox_tree_new(max(n, 1))
    batches, total, log = _insert_segments(t, metric, segments, grid_coords, grid_scale,
                                           use_grid)
    return _finish(t, metric, n, order, batches, total, log)


def _child_walk(children, nxt, u, n):
    """Children of u in list order (prepend lists, src/tree.rs:98-103)."""
    out = []
    c = int(children[u])
    first = c
    steps = 0
    while c != -1 and steps < n:
        out.append(c)
        nx = int(nxt[c])
        if nx == c or nx == first:
            break
        c = nx
        steps += 1
    return out


def delete_rows(tree: "OracleTree", data, remove, *, v=None, p=None, var=1.0, ls=1.0,
                kernel_type=0, chunk=512, grid_scale=1.0, exp_mode=EXP_LIBM,
                euclidean=False) -> "OracleTree":
    """batch_delete's procedure (covertreex/algo/batch_delete.py:105-525) on a
    tree of this oracle: removed nodes and all their descendants leave the
    tree; the remaining links stay; the detached (not removed) nodes go back
    in groups of equal max(level, 0), lowest first, each through chunked
    batch inserts.  ``data`` (and v, p for a residual tree) are the tree's
    dataset rows; ``remove`` holds dataset rows (residual) or node indices
    (Euclidean).  Remaining dataset rows are renumbered in order; the new tree
    order is [kept nodes, detached nodes by group]."""
    order, inv = tree.order, tree.inv_order
    parents, children, nxt, levels = tree.parents, tree.children, tree.next, tree.levels
    n = parents.size
    state = np.zeros(n, np.int8)
    removed_ds = np.zeros(n, bool)
    stack = []
    for x in np.asarray(remove, np.int64).reshape(-1):
        if x < 0 or x >= n:
            raise ValueError("Deletion index out of range.")
        r = int(order[x]) if euclidean else int(x)
        if removed_ds[r]:
            raise ValueError("Duplicate deletion indices are not allowed.")
        removed_ds[r] = True
        state[inv[r]] = 1
        stack.append(int(inv[r]))
    while stack:
        u = stack.pop()
        for c in _child_walk(children, nxt, u, n):
            if state[c] == 0:
                state[c] = 2
                stack.append(c)
    kept = [u for u in range(n) if state[u] == 0]
    glevel = lambda u: max(int(levels[u]), 0)  # noqa: E731
    det = sorted((u for u in range(n) if state[u] == 2), key=glevel)  # stable: tree order
    src = np.asarray(kept + det, np.int64)
    newpos = -np.ones(n, np.int64)
    newpos[src] = np.arange(src.size)
    n1, k0 = src.size, len(kept)
    ends = [k0 + i + 1 for i in range(len(det)) if i + 1 == len(det)
            or glevel(det[i + 1]) != glevel(det[i])]
    below = np.concatenate([[0], np.cumsum(removed_ds)])
    order1 = order[src] - below[order[src]]
    par1 = np.full(n1, -1, np.int64)
    chi1 = np.full(n1, -1, np.int64)
    nxt1 = np.full(n1, -1, np.int64)
    lev1 = np.full(n1, np.iinfo(np.int32).min, np.int32)
    for i, u in enumerate(kept):
        par1[i] = newpos[parents[u]] if parents[u] >= 0 else -1
        lev1[i] = levels[u]
        kids = [newpos[c] for c in _child_walk(children, nxt, u, n) if state[c] == 0]
        if kids:
            chi1[i] = kids[0]
            for a, b in zip(kids[:-1], kids[1:]):
                nxt1[a] = b
    # dataset rows in the new tree order (the old tree order, then src)
    rows = np.asarray(data, np.float32)[order][src]
    if euclidean:
        metric = euclidean_metric(rows)
        gc, use_grid = None, 0
    else:
        metric = residual_metric(rows, np.asarray(v, np.float32)[order][src],
                                 np.asarray(p, np.float32)[order][src], var, ls, kernel_type,
                                 exp_mode)
        gc, use_grid = rows, 1
    L = lib()
    t = L.ox_tree_new(max(n1, 1))
    tt = t.contents
    for dst, arr in ((tt.parents, par1), (tt.children, chi1), (tt.next, nxt1)):
        C.memmove(dst, np.ascontiguousarray(arr).ctypes.data, 8 * n1)
    C.memmove(tt.levels, np.ascontiguousarray(lev1).ctypes.data, 4 * n1)
    # the tree keeps its level range: bottom survivors take min_level (mod.rs:791-800)
    tt.min_level, tt.max_level = tree.min_level, tree.max_level
    tt.n = k0
    segs, g0 = [], k0
    for g1 in ends:
        segs.append((g0, g1, chunk))
        g0 = g1
    batches, total, log = _insert_segments(t, metric, segs, gc, grid_scale, use_grid)
    tt.n = n1
    return _finish(t, metric, n1, order1, batches, total, log)


def build_residual_tree(coords, v, p, *, var=1.0, ls=1.0, kernel_type=0, chunk=512,
                        batch_order="hilbert", grid_scale=1.0, exp_mode=EXP_LIBM) -> OracleTree:
    """build_pcct2_residual_tree (src/lib.rs:1373-1564) + compute_subtree_bounds_py."""
    coords = _c(coords, np.float32)
    n = coords.shape[0]
    if batch_order == "natural":
        order = np.arange(n, dtype=np.int64)
    else:
        order, _ = morton_order(coords)
    c_ord = coords[order]
    metric = residual_metric(c_ord, np.asarray(v, np.float32)[order], np.asarray(p, np.float32)[order],
                             var, ls, kernel_type, exp_mode)
    return _build(metric, n, order, [(0, n, chunk)], c_ord, grid_scale, 1)


def build_residual_tree_with_insert(coords, v, p, coords_new, v_new, p_new, *, var=1.0, ls=1.0,
                                    kernel_type=0, chunk=512, insert_chunk=None,
                                    batch_order="hilbert", grid_scale=1.0,
                                    exp_mode=EXP_LIBM) -> OracleTree:
    """The build above, then CoverTreeWrapper.insert_residual (src/lib.rs:254-300)
    of new rows: they become dataset rows n.. and tree nodes n.. in their
    given order, batch-inserted in chunks of insert_chunk (default: one batch)."""
    coords = _c(coords, np.float32)
    cn = _c(coords_new, np.float32)
    n0, n1 = coords.shape[0], coords.shape[0] + cn.shape[0]
    order0 = (np.arange(n0, dtype=np.int64) if batch_order == "natural"
              else morton_order(coords)[0])
    order = np.concatenate([order0, np.arange(n0, n1, dtype=np.int64)])
    c_all = np.concatenate([coords, cn])
    v_all = np.concatenate([np.asarray(v, np.float32), np.asarray(v_new, np.float32)])
    p_all = np.concatenate([np.asarray(p, np.float32), np.asarray(p_new, np.float32)])
    c_ord = c_all[order]
    metric = residual_metric(c_ord, v_all[order], p_all[order], var, ls, kernel_type, exp_mode)
    segs = [(0, n0, chunk), (n0, n1, insert_chunk or (n1 - n0))]
    return _build(metric, n1, order, segs, c_ord, grid_scale, 1)


def build_euclidean_tree(points, *, chunk=512, grid_scale=None) -> OracleTree:
    """Chunked batch_insert_collect with the Euclidean metric in natural order.

    With ``grid_scale`` the batch insert also adds grid conflict edges over
    the points themselves (``coords`` of batch_insert_collect,
    src/algo/batch/mod.rs:660-668), as the residual build does over its raw
    coordinates; the reference's Euclidean API path passes no coordinates."""
    pts = _c(points, np.float32)
    n = pts.shape[0]
    order = np.arange(n, dtype=np.int64)
    metric = euclidean_metric(pts)
    if grid_scale is None:
        return _build(metric, n, order, [(0, n, chunk)], None, 1.0, 0)
    return _build(metric, n, order, [(0, n, chunk)], pts, float(grid_scale), 1)


def build_euclidean_tree_with_insert(points, new_points, *, chunk=512,
                                     insert_chunk=None) -> OracleTree:
    """The Euclidean build, then CoverTreeWrapper.insert (src/lib.rs:207-223) of
    new points as nodes n.., in chunks of insert_chunk (default: one batch)."""
    pts = np.concatenate([_c(points, np.float32), _c(new_points, np.float32)])
    n0, n1 = np.asarray(points).shape[0], pts.shape[0]
    order = np.arange(n1, dtype=np.int64)
    metric = euclidean_metric(pts)
    return _build(metric, n1, order, [(0, n0, chunk), (n0, n1, insert_chunk or (n1 - n0))],
                  None, 1.0, 0)


def si_cache(tree: OracleTree) -> np.ndarray:
    si = np.zeros(max(tree.parents.size, 1), np.float32)
    lib().ox_si_cache(tree._tree, tree._metric.h, _ptr(si))
    return si[: tree.parents.size]


def tree_from_arrays(coords, v, p, *, parents, children, next_node, levels, si, var=1.0, ls=1.0,
                     kernel_type=0, exp_mode=EXP_LIBM) -> OracleTree:
    """A residual tree given node by node (natural order: node i is dataset
    row i), for hand-built query cases: flat arrays as CoverTreeData keeps
    them (src/tree.rs:6-125), si the cover radii the query reads."""
    coords = _c(coords, np.float32)
    n = coords.shape[0]
    metric = residual_metric(coords, np.asarray(v, np.float32), np.asarray(p, np.float32), var,
                             ls, kernel_type, exp_mode)
    L = lib()
    t = L.ox_tree_new(max(n, 1))
    tt = t.contents
    for name, arr, dt in (("parents", parents, np.int64), ("children", children, np.int64),
                          ("next", next_node, np.int64), ("levels", levels, np.int32),
                          ("si", si, np.float32)):
        a = _c(arr, dt)
        assert a.shape == (n,)
        C.memmove(getattr(tt, name), a.ctypes.data, a.nbytes)
    tt.n = n
    lv = np.asarray(levels, np.int64)
    set_lv = lv[lv != np.iinfo(np.int32).min]
    tt.min_level = int(set_lv.min()) if set_lv.size else 0
    tt.max_level = int(set_lv.max()) if set_lv.size else 0
    order = np.arange(n, dtype=np.int64)
    par = _c(parents, np.int64)
    smin, smax = subtree_bounds(par, order)
    return OracleTree(order=order, inv_order=order.copy(), parents=par.copy(),
                      children=_c(children, np.int64).copy(), next=_c(next_node, np.int64).copy(),
                      levels=_c(levels, np.int32).copy(), si=_c(si, np.float32).copy(),
                      min_level=tt.min_level, max_level=tt.max_level, subtree_min=smin,
                      subtree_max=smax, _tree=t, _metric=metric)


def _opts(budget, threads, budget_up=None, budget_down=None):
    o = _QueryOpts()
    lib().ox_query_opts_default(C.byref(o))
    if budget is not None and budget != "default":
        sched = list(budget)
        o.n_budget = len(sched)
        for i, b in enumerate(sched[:8]):
            o.budget[i] = int(b)
    if budget_up is not None:
        o.budget_up = float(budget_up)
    if budget_down is not None:
        o.budget_down = float(budget_down)
    o.threads = int(threads)
    return o


def knn_residual(tree: OracleTree, queries, k: int, *, predecessor=False, budget=(),
                 threads: int = 0, use_bounds=True, budget_up=None, budget_down=None):
    """CoverTreeWrapper.knn_query_residual cached path (src/lib.rs:369-507).

    ``budget=()`` is the exact search (empty schedule); ``budget="default"``
    the reference default ladder [32,64,96].  ``budget_up`` / ``budget_down``
    replace the yield thresholds 0.015 / 0.002 (the reference reads them from
    COVERTREEX_RESIDUAL_BUDGET_UP / _DOWN, src/algo.rs:661-676)."""
    q = _c(np.asarray(queries).reshape(-1), np.int64)
    n = tree.order.size
    q_tree = np.where((q >= 0) & (q < n), tree.inv_order[np.clip(q, 0, max(n - 1, 0))], 0)
    q_tree = _c(q_tree, np.int64)
    nq = q.size
    out_i = np.empty((nq, k), np.int64)
    out_d = np.empty((nq, k), np.float32)
    evals = np.zeros(nq, np.int64)
    o = _opts(budget, threads, budget_up, budget_down)
    smin = tree.subtree_min if (predecessor and use_bounds) else None
    lib().ox_knn_residual(tree._tree, tree._metric.h, _ptr(tree.order), _ptr(q_tree), _ptr(q),
                          nq, int(k), int(bool(predecessor)), _ptr(smin), C.byref(o),
                          _ptr(out_i), _ptr(out_d), _ptr(evals))
    return out_i, out_d, evals


def knn_euclidean(tree: OracleTree, queries, k: int, *, predecessor=False, threads=0):
    """batch_knn_query / single_knn_query (src/algo.rs:179-323)."""
    qp = _c(queries, np.float32)
    nq = qp.shape[0]
    out_i = np.empty((nq, k), np.int64)
    out_d = np.empty((nq, k), np.float32)
    lib().ox_knn_euclidean(tree._tree, tree._metric.h, _ptr(qp), nq, int(k),
                           int(bool(predecessor)), _ptr(out_i), _ptr(out_d), int(threads))
    return out_i, out_d
