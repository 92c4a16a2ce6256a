"""Hilbert order against golden vectors written by the reference's own
``covertreex/algo/order/strategy.py`` (``tests/golden/make_hilbert_golden.py``):
codes and permutations for D in {2, 3, 5, 8, 21}, duplicate rows, coarse ties
and a constant axis.  Pins the curve's orientation and the code's bit order."""

import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hilbert_golden.npz")
CASES = ["d2", "d3", "d5", "d8", "d21", "d3_const_axis"]


def _golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("case", CASES)
def test_oracle_hilbert_matches_reference(oracle_lib, case):
    g = _golden()
    pts = g[f"{case}_points"]
    assert np.array_equal(oracle.hilbert_codes(pts), g[f"{case}_codes"])
    assert np.array_equal(oracle.hilbert_order(pts), g[f"{case}_perm"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_device_hilbert_matches_reference(hip, case):
    from covertreex_amd.order import hilbert_permutation

    g = _golden()
    assert np.array_equal(hilbert_permutation(g[f"{case}_points"]), g[f"{case}_perm"])
