"""Pin the oracle's residual metric (src/metric.rs:376-563) against closed forms,
special cases and an FP64 evaluation; pin its exp modes."""

import math
import struct

import numpy as np
import pytest

import oracle
from oracle import fp64
from oracle import lib as ol


def _f32(x):
    return np.float32(x)


def _metric(coords, v, p, kt, var=1.0, ls=1.0, mode=ol.EXP_LIBM):
    return ol.residual_metric(coords, v, p, var, ls, kt, mode)


def test_zero_factor_closed_forms(oracle_lib):
    # V = 0, p = var = 1: rho = k(x_i, x_j), so 1 - |rho| = 1 - k
    x = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 0]], np.float32)
    v = np.zeros((3, 8), np.float32)
    p = np.ones(3, np.float32)
    rbf = _metric(x, v, p, 0)
    got = ol.dist_sq_pairs(rbf, [0, 0], [1, 2])
    assert got[0] == pytest.approx(1.0 - math.exp(-1.0), abs=2e-7)   # d2 = 2
    assert got[1] == pytest.approx(1.0 - math.exp(-0.5), abs=2e-7)   # d2 = 1
    mat = _metric(x, v, p, 1)
    a = math.sqrt(5.0) * math.sqrt(2.0)
    want = 1.0 - (1.0 + a + a * a / 3.0) * math.exp(-a)
    assert ol.dist_sq_pairs(mat, [0], [1])[0] == pytest.approx(want, abs=3e-7)
    # the query-tile variant returns the square root
    assert ol.dist_batch_pairs(mat, [0], [1])[0] == pytest.approx(math.sqrt(want), abs=3e-7)


def test_variance_and_lengthscale_scale_the_kernel(oracle_lib):
    x = np.array([[2.0, 0.0], [0.0, 0.0]], np.float32)
    v = np.zeros((2, 8), np.float32)
    p = np.full(2, 2.0, np.float32)
    m = _metric(x, v, p, 0, var=2.0, ls=2.0)  # d2 = (2/2)^2 = 1, k = 2 e^-0.5, rho = k / 2
    assert ol.dist_sq_pairs(m, [0], [1])[0] == pytest.approx(1.0 - math.exp(-0.5), abs=2e-7)


def test_self_zero_symmetric_clamped(oracle_lib):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((64, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=16, seed=0, kernel_type=1)
    m = _metric(x, v, p, 1)
    i, j = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    a = ol.dist_sq_pairs(m, i.ravel(), j.ravel()).reshape(64, 64)
    assert np.all(np.diag(a) == 0)
    assert np.array_equal(a.view(np.uint32), a.T.view(np.uint32))
    assert np.all((a >= 0) & (a <= 1))
    b = ol.dist_batch_pairs(m, i.ravel(), j.ravel()).reshape(64, 64)
    assert np.all(np.diag(b) == 0) and np.all((b >= 0) & (b <= 1))
    # |rho| >= 1 clamps to distance 0: a V row opposite to the kernel value
    x2 = np.zeros((2, 3), np.float32)
    v2 = np.zeros((2, 8), np.float32)
    v2[0, 0], v2[1, 0] = 2.0, -2.0          # V.V = -4, k = 1, rho = 5 / 1 -> 1
    m2 = _metric(x2, v2, np.ones(2, np.float32), 0)
    assert ol.dist_sq_pairs(m2, [0], [1])[0] == 0.0


def test_matches_fp64_on_well_conditioned_pairs(oracle_lib):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((400, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=24, seed=0, kernel_type=1)
    m = _metric(x, v, p, 1)
    a = rng.integers(0, 400, 3000)
    b = rng.integers(0, 400, 3000)
    keep = (p[a] > 0.05) & (p[b] > 0.05) & (a != b)
    a, b = a[keep], b[keep]
    d64 = np.array([fp64.residual_dist_fp64(x, v, p, [i], [j], kernel_type=1)[0, 0]
                    for i, j in zip(a[:400], b[:400])])
    d32 = ol.dist_batch_pairs(m, a[:400], b[:400]).astype(np.float64)
    sq32 = np.sqrt(ol.dist_sq_pairs(m, a[:400], b[:400]).astype(np.float64))
    assert np.max(np.abs(d32 - d64)) < 2e-4
    assert np.max(np.abs(sq32 - d64)) < 2e-4


def _bits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


def test_exp_modes(oracle_lib):
    rng = np.random.default_rng(2)
    xs = -np.abs(rng.standard_normal(20000).astype(np.float32)) * 30.0
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    got = np.array([ol.expf(float(x), ol.EXP_CR) for x in xs], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # glibc 2.35 expf is not correctly rounded everywhere: one documented input
    x = np.float32(float.fromhex("-0x1.9424fcp-14"))
    assert _bits(ol.expf(float(x), ol.EXP_CR)) == _bits(float.fromhex("0x1.fff35ep-1"))
    assert _bits(ol.expf(float(x), ol.EXP_LIBM)) == _bits(float.fromhex("0x1.fff36p-1"))
