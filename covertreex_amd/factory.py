"""``cover_tree()`` factory with the reference's signature plus ``engine="hip"``.

Mirrors covertreex/api/factory.py:40-348.  Residual trees are built with the
``hip`` engine; ``BuiltCoverTree.knn(k, queries=None, ...)`` queries every
point when ``queries`` is None (factory.py:100-102).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np

from .engine import CoverTree as EngineCoverTree
from .engine import HipEngine, RuntimeConfig
from .kernels import Kernel
from .residual import ResidualBackend, backend_from_kernel


@dataclass
class BuiltCoverTree:
    _tree: EngineCoverTree
    _n_points: int
    _runtime: RuntimeConfig
    _is_residual: bool = True

    def knn(self, k: int, *, queries: np.ndarray | Sequence[int] | None = None,
            return_distances: bool = False, predecessor_mode: bool = False):
        if queries is None:
            query_data = np.arange(self._n_points, dtype=np.int64).reshape(-1, 1)
        else:
            query_data = np.asarray(queries)
            if query_data.ndim == 1:
                query_data = query_data.reshape(-1, 1)
        return self._tree.knn(query_data, k=k, return_distances=return_distances,
                              predecessor_mode=predecessor_mode, runtime=self._runtime)

    @property
    def n_points(self) -> int:
        return self._n_points

    @property
    def build_time(self) -> float | None:
        return self._tree.build_seconds

    @property
    def tree(self) -> EngineCoverTree:
        return self._tree

    def __repr__(self) -> str:
        metric = "residual" if self._is_residual else "euclidean"
        return f"BuiltCoverTree(n={self._n_points}, metric={metric!r}, engine='hip')"


def cover_tree(coords: np.ndarray, *, kernel: Kernel | None = None,
               v_matrix: np.ndarray | None = None, p_diag: np.ndarray | None = None,
               kernel_diag: np.ndarray | None = None, inducing_count: int = 512, seed: int = 42,
               chunk_size: int = 512, engine: str = "hip",
               precision: str = "mixed") -> BuiltCoverTree:
    """Build a residual-correlation cover tree on the GPU.

    ``precision="mixed"`` (default) answers queries with the certified
    mixed-precision engine (exact neighbours under the residual metric on the
    stored inputs); ``"reference-f32"`` reproduces the reference's f32 search.
    """
    if engine not in ("hip", "rust-hilbert", "hilbert"):
        raise ValueError(f"engine {engine!r} is not available here; use engine='hip'")
    coords_f32 = np.asarray(coords, dtype=np.float32)
    n_points = coords_f32.shape[0]
    if kernel is None and v_matrix is None:
        from .euclidean import euclidean_cover_tree

        return euclidean_cover_tree(coords_f32, chunk_size=chunk_size)
    if kernel is not None and v_matrix is not None:
        raise ValueError("Provide either 'kernel' or 'v_matrix', not both.")
    if v_matrix is not None:
        if p_diag is None:
            raise ValueError("p_diag is required when v_matrix is provided.")
        kd = (np.ones(n_points, np.float32) if kernel_diag is None
              else np.asarray(kernel_diag, np.float32))
        backend = ResidualBackend(v_matrix=v_matrix, p_diag=p_diag, kernel_diag=kd,
                                  kernel_points_f32=coords_f32, chunk_size=chunk_size)
        # the reference's Rust metric reads RBF, var 1, ls 1 here (factory.py:261-265)
        params = {"variance": 1.0, "lengthscale": 1.0, "kernel_type": 0}
    else:
        backend = backend_from_kernel(coords_f32, kernel, seed=seed,
                                      inducing_count=inducing_count, chunk_size=chunk_size)
        params = {"variance": kernel.variance, "lengthscale": kernel.lengthscale,
                  "kernel_type": kernel.kernel_type}
    runtime = RuntimeConfig(metric="residual_correlation", engine="hip", knn_precision=precision)
    tree = HipEngine().build(coords_f32, runtime=runtime, residual_backend=backend,
                             residual_params=params, compute_predecessor_bounds=True)
    return BuiltCoverTree(_tree=tree, _n_points=n_points, _runtime=runtime, _is_residual=True)


__all__ = ["cover_tree", "BuiltCoverTree"]
