"""Pin the oracle's orderings: Morton (src/pcct.rs:10-67) and the python path's
Skilling Hilbert code (covertreex/algo/order/strategy.py:38-99)."""

import numpy as np
import pytest

import oracle


def _interleave3(x, y, z):
    code = 0
    for b in range(21):
        code |= ((x >> b) & 1) << (3 * b)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b + 2)
    return code


def test_morton_code_is_bit_interleave(oracle_lib):
    # coordinates k / 2^21 with anchors 0 and 1 quantise to trunc(k * (1 - 2^-21))
    rng = np.random.default_rng(3)
    k = rng.integers(0, 8, size=(200, 3))
    pts = np.vstack([np.zeros((1, 3)), np.ones((1, 3)), k / 2.0**21]).astype(np.float32)
    order, codes = oracle.morton_order(pts)
    q = np.floor(pts.astype(np.float32) * np.float32(2097151.0)).astype(np.int64)
    want = np.array([_interleave3(*map(int, r)) for r in q], dtype=np.uint64)
    assert np.array_equal(codes, want[order])


def test_morton_corners_known_order(oracle_lib):
    corners = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    perm = np.random.default_rng(0).permutation(8)
    order, _ = oracle.morton_order(corners[perm])
    # Z-order over {0,1}^3 sorts by (z, y, x)
    assert np.array_equal(perm[order], np.arange(8))


def test_morton_is_stable_permutation(oracle_lib):
    rng = np.random.default_rng(1)
    base = rng.standard_normal((50, 3)).astype(np.float32)
    pts = np.vstack([base, base[:20], base[:5]])
    order, codes = oracle.morton_order(pts)
    assert np.array_equal(np.sort(order), np.arange(pts.shape[0]))
    assert np.all(codes[1:] >= codes[:-1])
    same = codes[1:] == codes[:-1]
    assert np.all(order[1:][same] > order[:-1][same])


@pytest.mark.parametrize("d", [1, 2, 5, 8])
def test_morton_identity_for_non_3d(oracle_lib, d):
    pts = np.random.default_rng(d).standard_normal((37, d)).astype(np.float32)
    order, _ = oracle.morton_order(pts)
    assert np.array_equal(order, np.arange(37))


@pytest.mark.parametrize("d,side", [(2, 16), (3, 8), (8, 2)])
def test_hilbert_block_contiguous_and_adjacent(oracle_lib, d, side):
    bits = max(1, min(16, 64 // d))
    top = (1 << bits) - 1
    cells = np.stack(np.meshgrid(*([np.arange(side)] * d), indexing="ij"), -1).reshape(-1, d)
    anchors = np.array([np.zeros(d), np.full(d, top)])
    pts = np.vstack([anchors, cells]).astype(np.float64)
    codes = oracle.hilbert_codes(pts)[2:]
    # the aligned block at the origin holds the first side^d curve positions
    assert sorted(codes.tolist()) == list(range(side**d))
    walk = cells[np.argsort(codes)]
    steps = np.abs(np.diff(walk, axis=0)).sum(axis=1)
    assert np.all(steps == 1)


def test_hilbert_order_is_stable_lexsort(oracle_lib):
    pts = np.random.default_rng(5).standard_normal((64, 3))
    pts = np.vstack([pts, pts[:10]])
    order = oracle.hilbert_order(pts)
    codes = oracle.hilbert_codes(pts)
    assert np.array_equal(np.sort(order), np.arange(pts.shape[0]))
    sc = codes[order]
    assert np.all(sc[1:] >= sc[:-1])
    same = sc[1:] == sc[:-1]
    assert np.all(order[1:][same] > order[:-1][same])
