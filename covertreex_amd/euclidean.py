"""Euclidean cover tree on the GPU: the reference's Rust Euclidean path
(``CoverTreeWrapper.insert`` + ``knn_query``, src/lib.rs:207-357,
src/algo.rs:179-323) behind ``cover_tree(coords)`` without a kernel."""

from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import check, ptr
from .native import HipTree

_MODES = {"reference": _ffi.CX_KNN_REFERENCE_F32, "exact": _ffi.CX_KNN_EXACT}


class EuclideanHipTree(HipTree):
    """Euclidean tree; node i is input row i (natural order)."""

    @classmethod
    def build(cls, points, *, chunk_size: int = 512) -> "EuclideanHipTree":
        x = np.ascontiguousarray(points, dtype=np.float32)
        if x.ndim != 2:
            raise TypeError("expected a float32 or float64 2D array")
        out = C.c_void_p()
        check(_ffi.load().cx_tree_build_euclidean(ptr(x), x.shape[0], x.shape[1],
                                                  int(chunk_size), C.byref(out)))
        t = cls(out.value, x.shape[0])
        t.dim = x.shape[1]
        t._after_insert()
        return t

    def insert(self, points, *, chunk_size: int | None = None) -> None:
        """CoverTreeWrapper.insert (src/lib.rs:207-223) on the device tree: the
        points become nodes n, n+1, ... in their given order."""
        x = np.ascontiguousarray(points, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.dim:
            raise TypeError(f"expected a float32 or float64 2D array with {self.dim} columns")
        chunk = int(chunk_size or max(x.shape[0], 1))
        if chunk > 1024:
            raise ValueError("chunk_size must be <= 1024 on this engine; pass chunk_size")
        with self._lock:
            check(_ffi.load().cx_tree_insert_euclidean(self._h, ptr(x), x.shape[0], chunk))
            self.n += x.shape[0]
        self._after_insert()

    def knn_points(self, queries, k: int, *, predecessor_mode: bool = False,
                   mode: str = "exact"):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        if q.shape[1] != self.dim:
            raise ValueError(f"queries must have {self.dim} columns")
        idx = np.empty((q.shape[0], int(k)), np.int64)
        dist = np.empty((q.shape[0], int(k)), np.float32)
        with self._lock:
            check(_ffi.load().cx_tree_knn_euclidean(self._h, ptr(q), q.shape[0], int(k),
                                                     int(bool(predecessor_mode)), _MODES[mode],
                                                     ptr(idx), ptr(dist)))
        return idx, dist


@dataclass
class EuclideanCoverTree:
    """``cover_tree(coords)`` result: ``knn(k, queries=points)``."""

    tree: EuclideanHipTree
    n_points: int
    coords: np.ndarray
    mode: str = "exact"

    def knn(self, k: int, *, queries=None, return_distances: bool = False,
            predecessor_mode: bool = False):
        from .engine import _format_knn_output

        # queries=None queries every input point by its coordinates (the
        # reference passes indices as 1-D coordinates here, factory.py:84-98)
        q = self.coords if queries is None else np.asarray(queries, dtype=np.float32)
        idx, dist = self.tree.knn_points(q, k, predecessor_mode=predecessor_mode,
                                         mode=self.mode)
        return _format_knn_output(idx, dist, return_distances=return_distances)

    def __repr__(self) -> str:
        return f"BuiltCoverTree(n={self.n_points}, metric='euclidean', engine='hip')"


def euclidean_cover_tree(coords, *, chunk_size: int = 512, mode: str = "exact"):
    x = np.ascontiguousarray(coords, dtype=np.float32)
    t = EuclideanHipTree.build(x, chunk_size=chunk_size)
    return EuclideanCoverTree(tree=t, n_points=x.shape[0], coords=x, mode=mode)
