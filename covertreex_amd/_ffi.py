"""ctypes binding of ``libcovertreex_hip.so`` (C-ABI in ``include/covertreex_hip.h``).

The library is required: there is no CPU fallback.  Loading fails with a clear
error when it has not been built (``python -c "import __graft_entry__ as g;
g.build()"``) and every entry point raises :class:`HipError` on a non-zero
status, carrying the library's thread-local error text.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcovertreex_hip.so")

CX_OK = 0
CX_ERR_INVALID = 1
CX_ERR_HIP = 2
CX_ERR_NO_DEVICE = 3
CX_ERR_UNSUPPORTED = 4

CX_ORDER_HILBERT = 0
CX_ORDER_NATURAL = 1

CX_KNN_REFERENCE_F32 = 0
CX_KNN_MIXED = 1
CX_KNN_REFERENCE_F32_BUDGETED = 2
CX_KNN_EXACT = 1

# per-batch build telemetry fields (include/covertreex_hip.h CX_BF_*)
CX_BATCH_FIELDS = 16
BATCH_FIELDS = ("start", "size", "leaders", "survivors", "orphans", "far_levels",
                "scan_evals", "mis_evals", "nonleader_evals", "grid_edges", "distance_edges",
                "scan_ns", "mis_ns", "nonleader_ns", "total_ns", "conflict_ns")
# query telemetry layout (CX_QT_*)
CX_QT_SIZE = 160
CX_QT_LEVELS = 64
QT_SCALARS = ("mode", "queries", "distance_evals", "prunes_lower_bound",
              "prunes_lower_bound_chunks", "predecessor_filtered", "subtrees_pruned",
              "budget_escalations", "budget_early_terminate", "listed_candidates",
              "hard_queries", "hard_pairs", "max_list", "n_levels")
CX_QT_FRONTIER = 14
CX_QT_EXPANDED = 78
CX_QT_BLOCKS = 142
CX_QT_HARD_LISTED = 145
CX_QT_RESCREENED = 146
# cover-tree invariant counts (CX_INV_*)
CX_INV_FIELDS = 8
INV_FIELDS = ("covering", "nesting", "separation", "forest", "nodes", "pairs", "batches")

_i64 = C.c_int64
_i32 = C.c_int32
_f32 = C.c_float
_p = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/covertreex_hip.h
SIGNATURES = {
    "cx_last_error": (C.c_char_p, []),
    "cx_device_count": (_i32, [_p]),
    "cx_tree_build_residual": (_i32, [_p, _p, _p, _i64, _i32, _i32, _f32, _p, _i32, _i32, _i32,
                                      _i32, _f32, _p]),
    "cx_tree_knn_residual": (_i32, [_p, _p, _i64, _i32, _i32, _i32, _p, _p, _p]),
    "cx_tree_knn_residual_device": (_i32, [_p, _p, _i64, _i32, _i32, _i32, _p, _p, _p]),
    "cx_tree_build_euclidean": (_i32, [_p, _i64, _i32, _i32, _p]),
    "cx_tree_knn_euclidean": (_i32, [_p, _p, _i64, _i32, _i32, _i32, _p, _p]),
    "cx_tree_info": (_i32, [_p, _p, _p, _p, _p, _p]),
    "cx_tree_get_parents": (_i32, [_p, _p]),
    "cx_tree_get_children": (_i32, [_p, _p]),
    "cx_tree_get_next": (_i32, [_p, _p]),
    "cx_tree_get_levels": (_i32, [_p, _p]),
    "cx_tree_get_si_cache": (_i32, [_p, _p]),
    "cx_tree_get_order": (_i32, [_p, _p]),
    "cx_tree_get_subtree_bounds": (_i32, [_p, _p, _p]),
    "cx_subtree_bounds": (_i32, [_p, _p, _i64, _p, _p]),
    "cx_morton_order": (_i32, [_p, _i64, _i32, _p]),
    "cx_hilbert_order": (_i32, [_p, _i64, _i32, _p]),
    "cx_hilbert_order_noalloc": (_i32, [_p, _i64, _i32, _p, _p]),
    "cx_luby_mis": (_i32, [_p, _p, _i64, C.c_uint32, _p, _p]),
    "cx_nystrom_residual": (_i32, [_p, _i64, _i32, _p, _i32, C.c_double, _p, _i32, _i32,
                                   C.c_double, _p, _p]),
    "cx_tree_set_profiling": (_i32, [_p, _i32]),
    "cx_tree_set_screen_segment": (_i32, [_p, _i32]),
    "cx_set_conflict_log": (_i32, [_i32]),
    "cx_tree_conflict_log": (_i32, [_p, _p, _i64, _p]),
    "cx_tree_get_screen_segment": (_i32, [_p, _p]),
    "cx_tree_last_profile": (_i32, [_p, _p, _i32, _p]),
    "cx_tree_insert_residual": (_i32, [_p, _p, _p, _p, _i64, _i32]),
    "cx_tree_insert_euclidean": (_i32, [_p, _p, _i64, _i32]),
    "cx_tree_delete": (_i32, [_p, _p, _i64]),
    "cx_tree_build_telemetry": (_i32, [_p, _p, _i64, _p]),
    "cx_tree_check_invariants": (_i32, [_p, _p]),
    "cx_tree_debug_set_node": (_i32, [_p, _i64, _i64, _i32]),
    "cx_expf_cr_range": (_i32, [C.c_uint32, _i64, _p]),
    "cx_tree_set_query_telemetry": (_i32, [_p, _i32]),
    "cx_tree_last_query_telemetry": (_i32, [_p, _p, _i32, _p]),
    "cx_tree_clear_last_query_telemetry": (_i32, [_p]),
    "cx_tree_free": (None, [_p]),
}


class HipError(RuntimeError):
    """A non-zero cx_status from the HIP library."""

    def __init__(self, status: int, message: str):
        super().__init__(f"covertreex_hip error {status}: {message}")
        self.status = status


class HipValueError(HipError, ValueError):
    """CX_ERR_INVALID: an argument the library rejected."""


_lib = None


def load(path: str | None = None):
    """Load the HIP library (once).  Raises ``ImportError`` when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: the HIP engine has no CPU fallback; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != CX_OK:
        msg = load().cx_last_error()
        cls = HipValueError if status == CX_ERR_INVALID else HipError
        raise cls(int(status), msg.decode() if msg else "")


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    check(load().cx_device_count(C.byref(n)))
    return int(n.value)
