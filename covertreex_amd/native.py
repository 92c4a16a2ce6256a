"""Python handle over a device-resident tree (``cx_tree*``).

Mirrors the reference's ``CoverTreeWrapper`` surface (src/lib.rs:63-904):
build via :meth:`HipTree.build_residual` (``build_pcct2_residual_tree``),
query via :meth:`HipTree.knn_residual` (``knn_query_residual``) and the
``get_*`` getters.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _ffi
from ._ffi import check, ptr

MODES = {"reference": _ffi.CX_KNN_REFERENCE_F32, "reference-f32": _ffi.CX_KNN_REFERENCE_F32,
         "reference-budgeted": _ffi.CX_KNN_REFERENCE_F32_BUDGETED,
         "mixed": _ffi.CX_KNN_MIXED}


def _f32c(a, ndim):
    arr = np.asarray(a)
    if arr.dtype.kind not in "fiu":
        raise TypeError(f"expected a float32 or float64 {ndim}D array")
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    if arr.ndim != ndim:
        raise TypeError(f"expected a float32 or float64 {ndim}D array")
    return arr


class CoverTreeInvariantError(RuntimeError):
    """A device tree failed its cover-tree invariant check."""


class HipTree:
    """A cover tree resident in HBM on the current HIP device."""

    def __init__(self, handle: int, n: int):
        self._h = C.c_void_p(handle)
        self.n = int(n)
        self._lock = threading.Lock()
        # check the cover-tree invariants on the device after every insert
        # call (and after the build, by the builders), raising on a violation
        self.check_after_insert = os.environ.get("COVERTREEX_HIP_CHECK_INVARIANTS", "0") not in (
            "", "0", "false", "False")

    # -- construction ---------------------------------------------------------
    @classmethod
    def build_residual(cls, v_matrix, p_diag, coords, rbf_var: float, rbf_ls, *,
                       chunk_size: int = 512, batch_order: str = "hilbert",
                       grid_whiten_scale: float = 1.0, kernel_type: int = 0) -> "HipTree":
        """Build on the device (build_pcct2_residual_tree, src/lib.rs:1373-1564)."""
        lib = _ffi.load()
        v = _f32c(v_matrix, 2)
        p = _f32c(p_diag, 1)
        x = _f32c(coords, 2)
        ls = np.ascontiguousarray(np.atleast_1d(np.asarray(rbf_ls, dtype=np.float32)))
        n, m = v.shape
        if p.shape[0] != n or x.shape[0] != n:
            raise ValueError("v_matrix, p_diag and coords must have the same number of rows")
        order = (batch_order or "hilbert").lower()
        if order == "natural":
            mode = _ffi.CX_ORDER_NATURAL
        else:  # "hilbert", "hilbert-morton" and anything else -> Morton (src/lib.rs:1404-1411)
            mode = _ffi.CX_ORDER_HILBERT
        out = C.c_void_p()
        check(lib.cx_tree_build_residual(ptr(v), ptr(p), ptr(x), n, m, x.shape[1],
                                         float(rbf_var), ptr(ls), ls.size, int(kernel_type),
                                         int(chunk_size or n or 1), mode,
                                         float(grid_whiten_scale), C.byref(out)))
        t = cls(out.value, n)
        t._after_insert()
        return t

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _ffi.load().cx_tree_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- queries --------------------------------------------------------------
    def knn_residual(self, query_indices, k: int, *, predecessor_mode: bool = False,
                     mode: str = "mixed", return_error: bool = False):
        """kNN for dataset-index queries; (Q,k) int64 indices and float32 distances,
        ascending, padded with -1 / FLT_MAX (src/lib.rs:907-938)."""
        q = np.ascontiguousarray(np.asarray(query_indices).reshape(-1), dtype=np.int64)
        k = int(k)
        if k < 1:
            raise ValueError("k must be >= 1")
        idx = np.empty((q.size, k), np.int64)
        dist = np.empty((q.size, k), np.float32)
        err = np.empty((q.size, k), np.float32) if return_error else None
        with self._lock:
            check(_ffi.load().cx_tree_knn_residual(self._h, ptr(q), q.size, k,
                                                    int(bool(predecessor_mode)), MODES[mode],
                                                    ptr(idx), ptr(dist), ptr(err)))
        if return_error:
            return idx, dist, err
        return idx, dist

    def knn_residual_device(self, d_queries: int, nq: int, k: int, d_idx: int, d_dist: int,
                            d_err: int | None = None, *, predecessor_mode=False, mode="mixed"):
        """Device-pointer variant (inputs/outputs already in HBM)."""
        with self._lock:
            check(_ffi.load().cx_tree_knn_residual_device(
                self._h, C.c_void_p(d_queries), int(nq), int(k), int(bool(predecessor_mode)),
                MODES[mode], C.c_void_p(d_idx), C.c_void_p(d_dist),
                C.c_void_p(d_err) if d_err else None))

    # -- telemetry ------------------------------------------------------------------
    PHASES = {"mixed": ["prep", "query_sort", "screen_pass1", "hist_tau", "screen_pass2", "rescore",
                        "hard"],
              "reference": ["traversal"], "reference-budgeted": ["traversal"]}

    def set_screen_segment(self, seg: int = 0) -> None:
        """Candidates per screen segment of the mixed engine (0: default).
        Results do not depend on it; it is the tile shape the tuner sets."""
        with self._lock:
            check(_ffi.load().cx_tree_set_screen_segment(self._h, int(seg)))

    def screen_segment(self) -> int:
        """The segment length the last mixed query used (0 before the first)."""
        out = C.c_int32(0)
        with self._lock:
            check(_ffi.load().cx_tree_get_screen_segment(self._h, C.byref(out)))
        return int(out.value)

    def set_profiling(self, enable: bool = True) -> None:
        check(_ffi.load().cx_tree_set_profiling(self._h, int(bool(enable))))

    def last_profile(self) -> list:
        out = np.zeros(8, np.float32)
        n = C.c_int32(0)
        check(_ffi.load().cx_tree_last_profile(self._h, ptr(out), out.size, C.byref(n)))
        return [float(x) for x in out[: n.value]]

    # -- query telemetry (src/lib.rs:357-367, src/telemetry.rs:1-19) ------------
    def set_query_telemetry(self, enable: bool = True) -> None:
        """The COVERTREEX_RUST_QUERY_TELEMETRY switch, per tree."""
        check(_ffi.load().cx_tree_set_query_telemetry(self._h, int(bool(enable))))

    def last_query_telemetry(self) -> dict | None:
        """Counters of the last kNN call, keyed like ResidualQueryTelemetry
        (frontier_levels / frontier_expanded are summed over the call's queries
        per level; block_sizes lists the tile sizes used, one entry per tile
        kind), plus the mixed engine's own counters; None when not recorded."""
        out = np.zeros(_ffi.CX_QT_SIZE, np.int64)
        n = C.c_int32(0)
        check(_ffi.load().cx_tree_last_query_telemetry(self._h, ptr(out), out.size, C.byref(n)))
        if n.value == 0:
            return None
        t = {k: int(out[i]) for i, k in enumerate(_ffi.QT_SCALARS)}
        mode = {v: k for k, v in MODES.items() if k != "reference"}.get(t.pop("mode"), "reference")
        lv = t.pop("n_levels")
        blocks = out[_ffi.CX_QT_BLOCKS:_ffi.CX_QT_BLOCKS + 3]
        t.update({
            "engine_mode": mode,
            "frontier_levels": [int(x) for x in out[_ffi.CX_QT_FRONTIER:_ffi.CX_QT_FRONTIER + lv]],
            "frontier_expanded": [int(x) for x in
                                  out[_ffi.CX_QT_EXPANDED:_ffi.CX_QT_EXPANDED + lv]],
            "yields": [],
            "caps_applied": 0,
            "prunes_cap": 0,
            "masked_dedup": 0,
            "level_cache_hits": 0,
            "level_cache_misses": 0,
            # tiles per size instead of the reference's one entry per tile
            "block_sizes": [],
            "block_size_counts": {size: int(c) for size, c in zip((16, 32, 64), blocks)},
            "hard_listed": int(out[_ffi.CX_QT_HARD_LISTED]),
            "rescreened_queries": int(out[_ffi.CX_QT_RESCREENED]),
        })
        return t

    def clear_last_query_telemetry(self) -> None:
        check(_ffi.load().cx_tree_clear_last_query_telemetry(self._h))

    # -- build telemetry (BatchInsertTelemetry, src/lib.rs:1274-1371) -------------
    def build_telemetry(self) -> list:
        """One dict per batch of the build and of later inserts."""
        nb = C.c_int64(0)
        L = _ffi.load()
        check(L.cx_tree_build_telemetry(self._h, None, 0, C.byref(nb)))
        out = np.zeros((max(nb.value, 1), _ffi.CX_BATCH_FIELDS), np.int64)
        check(L.cx_tree_build_telemetry(self._h, ptr(out), nb.value, C.byref(nb)))
        return [{k: int(row[i]) for i, k in enumerate(_ffi.BATCH_FIELDS)}
                for row in out[: nb.value]]

    def conflict_log(self) -> np.ndarray:
        """The conflict records of the batches inserted while
        :func:`set_conflict_log` was on (int32 words; layout in
        ``include/covertreex_hip.h``; ``telemetry.batch_payloads`` turns them
        into the reference's per-batch payload)."""
        nw = C.c_int64(0)
        L = _ffi.load()
        check(L.cx_tree_conflict_log(self._h, None, 0, C.byref(nw)))
        out = np.zeros(max(nw.value, 1), np.int32)
        check(L.cx_tree_conflict_log(self._h, ptr(out), nw.value, C.byref(nw)))
        return out[: nw.value]

    # -- incremental insert (src/lib.rs:254-300) ----------------------------------
    def insert_residual(self, v_new, p_new, coords_new, *, chunk_size: int | None = None):
        """Append rows (dataset indices n, n+1, ...) and batch-insert them in
        their given order; chunk_size defaults to one batch of all the rows
        (the reference's default), at most 1024."""
        v = np.ascontiguousarray(v_new, np.float32)
        p = np.ascontiguousarray(p_new, np.float32).reshape(-1)
        x = np.ascontiguousarray(coords_new, np.float32)
        if v.ndim != 2 or x.ndim != 2 or v.shape[0] != p.size or x.shape[0] != p.size:
            raise TypeError("expected V (n, m), p (n,) and coords (n, d) float arrays")
        chunk = int(chunk_size or max(p.size, 1))
        if chunk > 1024:
            raise ValueError("chunk_size must be <= 1024 on this engine; pass chunk_size")
        with self._lock:
            check(_ffi.load().cx_tree_insert_residual(self._h, ptr(v), ptr(p), ptr(x), p.size,
                                                      chunk))
            self.n += p.size
        self._after_insert()

    # -- delete (covertreex/algo/batch_delete.py:628-689) --------------------------
    def delete(self, indices) -> None:
        """Remove points (dataset rows of a residual tree; node indices of a
        Euclidean tree): their descendants are re-inserted, remaining rows are
        renumbered in order.  ValueError on duplicates or out-of-range rows."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int64).reshape(-1))
        with self._lock:
            check(_ffi.load().cx_tree_delete(self._h, ptr(idx), idx.size))
            self.n -= idx.size
        self._after_insert()

    # -- cover-tree invariants ----------------------------------------------------
    def check_invariants(self, *, raise_on_violation: bool = True) -> dict:
        """Covering (d(node, parent) <= 2^(level+1)), nesting (a node's parent
        was in the tree when it was inserted, and the parent chain reaches a
        root), separation (nodes one batch inserted at one level
        pairwise farther than 2^level and, with grid conflicts, not grid
        neighbours) and forest consistency, checked on the device: violation
        counts plus what was checked (nodes, pairs, batches).  These are the
        properties the reference's batch insert guarantees
        (src/algo/batch/mod.rs:447-800); separation covers the batches
        inserted since the build or the last delete.  Raises
        CoverTreeInvariantError on a violation when raise_on_violation."""
        out = np.zeros(_ffi.CX_INV_FIELDS, np.int64)
        with self._lock:
            check(_ffi.load().cx_tree_check_invariants(self._h, ptr(out)))
        res = {k: int(out[i]) for i, k in enumerate(_ffi.INV_FIELDS)}
        bad = {k: res[k] for k in ("covering", "nesting", "separation", "forest") if res[k]}
        if bad and raise_on_violation:
            raise CoverTreeInvariantError(f"cover-tree invariants violated: {bad} ({res})")
        return res

    def debug_set_node(self, node: int, parent: int, level: int) -> None:
        """Test support: overwrite one node's parent pointer and level on the
        device without relinking (cx_tree_debug_set_node)."""
        check(_ffi.load().cx_tree_debug_set_node(self._h, int(node), int(parent), int(level)))

    def _after_insert(self) -> None:
        if self.check_after_insert:
            self.check_invariants()

    # -- getters (src/lib.rs:825-904) -------------------------------------------
    def info(self):
        n, m, d = C.c_int64(), C.c_int32(), C.c_int32()
        lo, hi = C.c_int32(), C.c_int32()
        check(_ffi.load().cx_tree_info(self._h, C.byref(n), C.byref(m), C.byref(d), C.byref(lo),
                                       C.byref(hi)))
        return {"n": n.value, "m": m.value, "d": d.value, "min_level": lo.value,
                "max_level": hi.value}

    def _get(self, fn, dtype):
        out = np.empty(self.n, dtype)
        check(getattr(_ffi.load(), fn)(self._h, ptr(out)))
        return out

    def get_parents(self):
        return self._get("cx_tree_get_parents", np.int64)

    def get_children(self):
        return self._get("cx_tree_get_children", np.int64)

    def get_next_node(self):
        return self._get("cx_tree_get_next", np.int64)

    def get_levels(self):
        return self._get("cx_tree_get_levels", np.int32)

    def get_si_cache(self):
        return self._get("cx_tree_get_si_cache", np.float32)

    def get_order(self):
        return self._get("cx_tree_get_order", np.int64)

    def get_min_level(self):
        return self.info()["min_level"]

    def get_max_level(self):
        return self.info()["max_level"]

    def get_subtree_bounds(self):
        mn = np.empty(self.n, np.int64)
        mx = np.empty(self.n, np.int64)
        check(_ffi.load().cx_tree_get_subtree_bounds(self._h, ptr(mn), ptr(mx)))
        return mn, mx


def set_conflict_log(enable: bool = True) -> None:
    """Record the conflict graph of every later build and insert in this
    process (the reference collects it when a log writer is given,
    src/algo/batch/mod.rs:802-956); off by default."""
    check(_ffi.load().cx_set_conflict_log(int(bool(enable))))


def compute_subtree_bounds(parents, node_to_dataset):
    """compute_subtree_bounds_py (src/lib.rs:991-1006) on the device."""
    par = np.ascontiguousarray(parents, dtype=np.int64)
    n2d = np.ascontiguousarray(node_to_dataset, dtype=np.int64)
    if par.shape != n2d.shape:
        raise ValueError("parents and node_to_dataset must have the same length")
    mn = np.empty(par.size, np.int64)
    mx = np.empty(par.size, np.int64)
    check(_ffi.load().cx_subtree_bounds(ptr(par), ptr(n2d), par.size, ptr(mn), ptr(mx)))
    return mn, mx


def hilbert_order(coords):
    """hilbert_order (src/lib.rs:956-962): Morton permutation for D == 3."""
    x = _f32c(coords, 2)
    out = np.empty(x.shape[0], np.int64)
    check(_ffi.load().cx_morton_order(ptr(x), x.shape[0], x.shape[1], ptr(out)))
    return out
