"""Nystrom factor on the GPU (cx_nystrom_residual)."""

from __future__ import annotations

import numpy as np

from . import _ffi
from ._ffi import check, ptr


def nystrom_device(points, inducing_idx, *, variance: float = 1.0, lengthscale=1.0,
                   kernel_type: int = 0, jitter: float = 1e-6):
    """V (N, m) float32 and p (N,) float32 for sorted inducing indices.

    Follows build_residual_backend (covertreex/metrics/residual/host_backend.py:201-240)
    in FP64 on the device."""
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if pts.ndim != 2:
        raise ValueError("points must be (N, D)")
    ind = np.ascontiguousarray(inducing_idx, dtype=np.int64)
    ls = np.ascontiguousarray(np.atleast_1d(np.asarray(lengthscale, dtype=np.float64)))
    n, d = pts.shape
    m = ind.size
    v = np.empty((n, m), np.float32)
    p = np.empty(n, np.float32)
    check(_ffi.load().cx_nystrom_residual(ptr(pts), n, d, ptr(ind), m, float(variance), ptr(ls),
                                          ls.size, int(kernel_type), float(jitter), ptr(v),
                                          ptr(p)))
    return v, p
