"""Python-path build primitives on the GPU: Hilbert batch order and Luby MIS
(covertreex/algo/order/strategy.py:38-99, covertreex/algo/_mis_numba.py:19-131)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check, ptr


def hilbert_permutation(points) -> np.ndarray:
    """Stable Skilling-Hilbert permutation of (N, D) points (D <= 64)."""
    x = np.ascontiguousarray(points, dtype=np.float64)
    if x.ndim != 2:
        raise ValueError("Expected a 2-D array of points for batch ordering.")
    out = np.empty(x.shape[0], np.int64)
    check(_ffi.load().cx_hilbert_order(ptr(x), x.shape[0], x.shape[1], ptr(out)))
    return out


def run_mis(indptr, indices, seed: int) -> tuple[np.ndarray, int]:
    """Luby MIS of a CSR graph: (int8 indicator, iterations)."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int64)
    n = ip.size - 1
    sel = np.zeros(max(n, 1), np.uint8)
    it = C.c_int64(0)
    check(_ffi.load().cx_luby_mis(ptr(ip), ptr(ix) if ix.size else None, n,
                                  int(seed) & 0xFFFFFFFF, ptr(sel), C.byref(it)))
    return sel[:n].astype(np.int8), int(it.value)
