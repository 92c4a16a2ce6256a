"""GP kernels for the residual-correlation metric (covertreex/kernels.py:59-132).

``kernel_type`` 0 = RBF, 1 = Matern 5/2, as the reference's Rust metric reads it
(src/metric.rs:44-57).  ``lengthscale`` may be a scalar or a per-dimension
vector (ARD); unlike the reference's rust-hilbert engine, which ignores it
(covertreex/engine.py:1002-1010), the HIP engine uses it in the metric.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Union

import numpy as np

LengthScale = Union[float, tuple, np.ndarray]


def _scaled_sqdist(x1, x2, lengthscale):
    ls = np.maximum(np.asarray(lengthscale, dtype=np.float64), 1e-12)
    a = np.asarray(x1, np.float64) / ls
    b = np.asarray(x2, np.float64) / ls
    diff = a[:, None, :] - b[None, :, :]
    return np.sum(diff * diff, axis=2)


class Kernel(ABC):
    variance: float
    lengthscale: LengthScale

    @abstractmethod
    def __call__(self, X1: np.ndarray, X2: np.ndarray) -> np.ndarray: ...

    @property
    def kernel_type(self) -> int:
        return 0


@dataclass(frozen=True)
class RBF(Kernel):
    """K(x, y) = variance * exp(-0.5 * ||x - y||^2 / lengthscale^2)."""

    lengthscale: LengthScale = 1.0
    variance: float = 1.0

    def __call__(self, X1, X2):
        return float(self.variance) * np.exp(-0.5 * _scaled_sqdist(X1, X2, self.lengthscale))

    @property
    def kernel_type(self) -> int:
        return 0


@dataclass(frozen=True)
class Matern52(Kernel):
    """K(x, y) = variance * (1 + a + a^2/3) * exp(-a), a = sqrt(5) ||x - y|| / lengthscale."""

    lengthscale: LengthScale = 1.0
    variance: float = 1.0

    def __call__(self, X1, X2):
        a = np.sqrt(5.0) * np.sqrt(_scaled_sqdist(X1, X2, self.lengthscale))
        return float(self.variance) * (1.0 + a + a * a / 3.0) * np.exp(-a)

    @property
    def kernel_type(self) -> int:
        return 1


__all__ = ["Kernel", "RBF", "Matern52"]
