"""Residual-correlation backend: Nystrom factor V, p_diag and kernel parameters.

Mirrors ``ResidualCorrHostData`` (covertreex/metrics/residual/core.py:45-154) and
``build_residual_backend`` (covertreex/metrics/residual/host_backend.py:165-246):
inducing indices ``default_rng(seed).choice(N, m, replace=False)`` sorted,
``K_mm + 1e-6 * var * I``, FP64 Cholesky, ``V = L^-1 K_xm^T`` transposed,
``p = max(var - |V|^2, 1e-9)`` in FP64, then FP32 storage.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .kernels import Kernel, Matern52, RBF


@dataclass(frozen=True)
class ResidualBackend:
    v_matrix: np.ndarray          # (N, m) float32
    p_diag: np.ndarray            # (N,) float32
    kernel_diag: np.ndarray       # (N,) float32
    kernel_points_f32: np.ndarray  # (N, D) float32
    variance: float = 1.0
    lengthscale: np.ndarray | float = 1.0
    kernel_type: int = 0
    chunk_size: int = 512
    inducing_idx: np.ndarray | None = None
    extras: dict = field(default_factory=dict)

    def __post_init__(self):
        v = np.ascontiguousarray(self.v_matrix, dtype=np.float32)
        p = np.ascontiguousarray(self.p_diag, dtype=np.float32)
        kd = np.ascontiguousarray(self.kernel_diag, dtype=np.float32)
        x = np.ascontiguousarray(self.kernel_points_f32, dtype=np.float32)
        if v.ndim != 2:
            raise ValueError("v_matrix must be two-dimensional.")
        if p.ndim != 1:
            raise ValueError("p_diag must be one-dimensional.")
        if v.shape[0] != p.shape[0]:
            raise ValueError("v_matrix and p_diag must have consistent length.")
        if kd.shape[0] != p.shape[0]:
            raise ValueError("kernel_diag must align with the cached points.")
        if x.ndim != 2 or x.shape[0] != p.shape[0]:
            raise ValueError("kernel_points_f32 must be (N, D) aligned with p_diag.")
        object.__setattr__(self, "v_matrix", v)
        object.__setattr__(self, "p_diag", p)
        object.__setattr__(self, "kernel_diag", kd)
        object.__setattr__(self, "kernel_points_f32", x)

    @property
    def num_points(self) -> int:
        return int(self.v_matrix.shape[0])

    @property
    def rank(self) -> int:
        return int(self.v_matrix.shape[1])


def inducing_indices(n_points: int, inducing_count: int, seed: int) -> np.ndarray:
    """host_backend.py:191-199."""
    m = min(int(inducing_count), n_points)
    if m <= 0:
        m = min(32, n_points)
    if m < n_points:
        return np.sort(np.random.default_rng(seed).choice(n_points, size=m, replace=False))
    return np.arange(n_points)


def build_residual_backend(points, *, seed: int, inducing_count: int, variance: float = 1.0,
                           lengthscale=1.0, chunk_size: int = 512,
                           kernel_type: int = 0) -> ResidualBackend:
    """Nystrom factor on the GPU (FP64 kernel matrices, Cholesky, triangular
    solve and p_diag; see ``covertreex_amd.nystrom``)."""
    from .nystrom import nystrom_device

    pts = np.asarray(points)
    if pts.size == 0:
        raise ValueError("Residual metric requires at least one point to configure caches.")
    pts32 = np.ascontiguousarray(pts, dtype=np.float32)
    idx = inducing_indices(pts.shape[0], inducing_count, seed)
    v, p = nystrom_device(np.asarray(points, dtype=np.float64), idx, variance=float(variance),
                          lengthscale=lengthscale, kernel_type=int(kernel_type))
    kdiag = np.full(pts.shape[0], variance, dtype=np.float32)
    return ResidualBackend(v_matrix=v, p_diag=p, kernel_diag=kdiag, kernel_points_f32=pts32,
                           variance=float(variance), lengthscale=lengthscale,
                           kernel_type=int(kernel_type), chunk_size=int(chunk_size),
                           inducing_idx=idx)


def backend_from_kernel(points, kernel: Kernel, *, seed: int, inducing_count: int,
                        chunk_size: int = 512) -> ResidualBackend:
    kt = kernel.kernel_type if isinstance(kernel, (RBF, Matern52)) else int(
        getattr(kernel, "kernel_type", 0))
    return build_residual_backend(points, seed=seed, inducing_count=inducing_count,
                                  variance=kernel.variance, lengthscale=kernel.lengthscale,
                                  chunk_size=chunk_size, kernel_type=kt)


__all__ = ["ResidualBackend", "build_residual_backend", "backend_from_kernel",
           "inducing_indices"]
