"""``Runtime`` / ``CoverTree`` API surface (covertreex/api/runtime.py:146-371,
covertreex/api/pcct.py:55-305) routed to the hip engine.

Unlike the reference, whose ``CoverTree.fit`` ignores ``engine`` and dispatches
on ``enable_rust`` (covertreex/algo/batch/insert.py:150-162), ``fit`` here always
builds on the GPU: residual trees with :class:`HipEngine`, Euclidean trees with
the device Euclidean path.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np

from .engine import HipEngine, RuntimeConfig
from .residual import ResidualBackend, build_residual_backend


@dataclass(frozen=True)
class Residual:
    """Residual-correlation parameters (covertreex/api/runtime.py Residual)."""

    variance: float = 1.0
    lengthscale: Any = 1.0
    kernel_type: int = 0
    inducing_count: int = 512
    seed: int = 42
    chunk_size: int = 512


@dataclass(frozen=True)
class Runtime:
    metric: str = "euclidean"
    engine: str = "hip"
    batch_order: str = "hilbert"
    grid_whiten_scale: float = 1.0
    knn_precision: str = "mixed"
    residual: Residual = field(default_factory=Residual)
    residual_backend: ResidualBackend | None = None

    def with_updates(self, **kw) -> "Runtime":
        return replace(self, **kw)

    def to_config(self) -> RuntimeConfig:
        return RuntimeConfig(metric="residual_correlation" if self.metric.startswith("residual")
                             else "euclidean", engine=self.engine,
                             batch_order_strategy=self.batch_order,
                             residual_grid_whiten_scale=self.grid_whiten_scale,
                             knn_precision=self.knn_precision)


class CoverTree:
    """Build with :meth:`fit`, query with :meth:`knn`."""

    def __init__(self, runtime: Runtime | None = None):
        self.runtime = runtime or Runtime()
        self._tree = None
        self._n = 0

    @property
    def is_residual(self) -> bool:
        return self.runtime.metric.startswith("residual")

    def fit(self, points, *, residual_backend: ResidualBackend | None = None) -> "CoverTree":
        pts = np.asarray(points, dtype=np.float32)
        if pts.ndim != 2:
            raise ValueError("points must be (N, D)")
        self._n = pts.shape[0]
        if self.is_residual:
            b = residual_backend or self.runtime.residual_backend
            r = self.runtime.residual
            if b is None:
                b = build_residual_backend(pts, seed=r.seed, inducing_count=r.inducing_count,
                                           variance=r.variance, lengthscale=r.lengthscale,
                                           chunk_size=r.chunk_size, kernel_type=r.kernel_type)
            params = {"variance": b.variance, "lengthscale": b.lengthscale,
                      "kernel_type": b.kernel_type, "chunk_size": r.chunk_size}
            self._tree = HipEngine().build(pts, runtime=self.runtime.to_config(),
                                           residual_backend=b, residual_params=params)
        else:
            from .euclidean import euclidean_cover_tree

            self._tree = euclidean_cover_tree(pts, chunk_size=self.runtime.residual.chunk_size)
        return self

    def knn(self, query_points, *, k: int, return_distances: bool = False,
            predecessor_mode: bool = False):
        if self._tree is None:
            raise RuntimeError("CoverTree.fit must be called before knn")
        if self.is_residual:
            return self._tree.knn(query_points, k=k, return_distances=return_distances,
                                  predecessor_mode=predecessor_mode,
                                  runtime=self.runtime.to_config())
        return self._tree.knn(k, queries=query_points, return_distances=return_distances,
                              predecessor_mode=predecessor_mode)

    @property
    def num_points(self) -> int:
        return self._n


__all__ = ["Runtime", "Residual", "CoverTree"]
