"""Device tree build and reference-form query vs the oracle (bit for bit).

The oracle runs with OX_EXP_CR, the double-evaluated expf the device kernels
use; the oracle's libm mode (the reference on this platform) is compared with
it in tests/test_oracle_tree.py."""

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu


def _data(n, d, m, kt, seed=0, dist="normal", ls=1.0):
    rng = np.random.default_rng(seed)
    if dist == "normal":
        x = rng.standard_normal((n, d)).astype(np.float32)
    else:
        x = rng.uniform(-3, 3, (n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt, lengthscale=ls)
    return x, v, p


BUILD_CASES = [
    # C1 shape: N=4096, D=3, Matern, m=64
    dict(n=4096, d=3, m=64, kt=1, chunk=512, gs=1.0),
    dict(n=2000, d=2, m=32, kt=0, chunk=64, gs=8.0),
    dict(n=3000, d=5, m=48, kt=0, chunk=256, gs=1.0, dist="uniform"),
    dict(n=1500, d=3, m=40, kt=1, chunk=1024, gs=4.0),
    dict(n=1, d=3, m=1, kt=1, chunk=512, gs=1.0),
    dict(n=513, d=3, m=16, kt=0, chunk=512, gs=1.0),
    # smooth field: p ~ 1e-7 for most rows, residual correlations at rounding
    # level, many candidates kept at their cover level after level (the
    # incremental-scan path) and a root with thousands of children
    dict(n=8000, d=2, m=256, kt=0, chunk=512, gs=1.0, ls=1.5),
    # ranks past 256 and between the powers of two: the cooperative dots of
    # the descent, wide scans and conflict pairs (64 chunks of eight at
    # m = 512), the fused levels' per-lane dots at m = 512
    dict(n=3000, d=5, m=512, kt=0, chunk=512, gs=1.0, dist="uniform"),
    dict(n=2500, d=3, m=200, kt=1, chunk=512, gs=1.0),
]


def _same_sets_modulo_cut_ties(hi, hd, oi, od):
    ties = 0
    for r in range(hi.shape[0]):
        if np.array_equal(hi[r], oi[r]):
            continue
        valid = hi[r] >= 0
        cut = hd[r][valid].max()
        below = hd[r] < cut
        assert set(hi[r][below & valid]) == set(oi[r][(od[r] < cut) & (oi[r] >= 0)])
        for dv in np.unique(hd[r][below & valid]):
            assert set(hi[r][hd[r] == dv]) == set(oi[r][od[r] == dv])
        ties += 1
    return ties


def _tree_pair(hip, case, seed=0):
    from covertreex_amd.native import HipTree

    ls = case.get("ls", 1.0)
    x, v, p = _data(case["n"], case["d"], case["m"], case["kt"], seed=seed,
                    dist=case.get("dist", "normal"), ls=ls)
    ot = oracle.build_residual_tree(x, v, p, ls=ls, kernel_type=case["kt"], chunk=case["chunk"],
                                    grid_scale=case["gs"], exp_mode=oracle.EXP_CR)
    ht = HipTree.build_residual(v, p, x, 1.0, np.full(1, ls, np.float32),
                                chunk_size=case["chunk"], grid_whiten_scale=case["gs"],
                                kernel_type=case["kt"])
    return x, v, p, ot, ht


@pytest.mark.parametrize("case", BUILD_CASES)
def test_device_tree_bit_exact(hip, case):
    print(f"case {case}", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    assert np.array_equal(ht.get_order(), ot.order)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    assert np.array_equal(ht.get_si_cache().view(np.uint32), ot.si.view(np.uint32))
    mn, mx = ht.get_subtree_bounds()
    assert np.array_equal(mn, ot.subtree_min) and np.array_equal(mx, ot.subtree_max)
    info = ht.info()
    assert (info["min_level"], info["max_level"]) == (ot.min_level, ot.max_level)


@pytest.mark.parametrize("case,k,pred", [
    (BUILD_CASES[0], 16, False),   # C1
    (BUILD_CASES[0], 16, True),
    (BUILD_CASES[1], 10, True),    # proper tree: pruning active
    (BUILD_CASES[1], 7, False),
    (BUILD_CASES[2], 30, True),
    (BUILD_CASES[3], 64, True),
])
def test_reference_form_knn_matches_oracle(hip, case, k, pred):
    print(f"case {case} k={k} pred={pred}", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    q = np.arange(case["n"])
    oi, od, _ = oracle.knn_residual(ot, q, k, predecessor=pred, budget=())
    hi, hd = ht.knn_residual(q, k, predecessor_mode=pred, mode="reference")
    # distances bit-exact; index sets equal except which members of an exact
    # f32 tie at the k-th distance survive the cut
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    _same_sets_modulo_cut_ties(hi, hd, oi, od)


@pytest.mark.parametrize("case,k,pred", [
    (BUILD_CASES[0], 16, False),
    (BUILD_CASES[1], 10, True),
    (BUILD_CASES[1], 10, False),
    (BUILD_CASES[2], 30, True),
])
def test_budgeted_reference_form_matches_oracle(hip, case, k, pred):
    # the reference's default search: budget ladder [32, 64, 96]
    print(f"case {case} k={k} pred={pred}", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    q = np.arange(case["n"])
    oi, od, _ = oracle.knn_residual(ot, q, k, predecessor=pred, budget="default")
    hi, hd = ht.knn_residual(q, k, predecessor_mode=pred, mode="reference-budgeted")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    _same_sets_modulo_cut_ties(hi, hd, oi, od)


def test_reference_form_edge_queries(hip):
    case = BUILD_CASES[1]
    x, v, p, ot, ht = _tree_pair(hip, case, seed=3)
    q = np.array([0, 1, 2, case["n"] - 1, 5, 5])
    oi, od, _ = oracle.knn_residual(ot, q, 12, predecessor=True, budget=())
    hi, hd = ht.knn_residual(q, 12, predecessor_mode=True, mode="reference")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    assert np.all(hi[0] == -1) and np.all(hd[0] == np.finfo(np.float32).max)
    assert set(hi[1][hi[1] >= 0]) == {0}


def test_c2_shape_build_and_reference_queries(hip):
    # N=32768, D=3, Matern, m=256: full build bit-exact, a query sample bit-exact
    case = dict(n=32768, d=3, m=256, kt=1, chunk=512, gs=1.0)
    print("C2 shape", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    assert np.array_equal(ht.get_si_cache().view(np.uint32), ot.si.view(np.uint32))
    q = np.random.default_rng(7).choice(32768, 256, replace=False)
    oi, od, _ = oracle.knn_residual(ot, q, 50, predecessor=True, budget=())
    hi, hd = ht.knn_residual(q, 50, predecessor_mode=True, mode="reference")
    assert np.array_equal(hd.view(np.uint32), od.view(np.uint32))
    _same_sets_modulo_cut_ties(hi, hd, oi, od)


@pytest.mark.parametrize("fused", ["-11", "-1", "-5"])
@pytest.mark.parametrize("case", [BUILD_CASES[0], BUILD_CASES[1], BUILD_CASES[6]])
def test_build_split_between_wide_and_fused_levels(hip, monkeypatch, case, fused):
    # where the batch loop switches from three launches per level to the
    # one-workgroup kernel does not change the tree (-11: no level fused,
    # -1: every level below the descent fused)
    monkeypatch.setenv("COVERTREEX_HIP_FUSED_LEVEL", fused)
    print(f"case {case} fused from {fused}", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    tel = ht.build_telemetry()
    assert len(tel) == len(ot.batch_stats)
    for h, o in zip(tel, ot.batch_stats):
        assert (h["scan_evals"], h["mis_evals"], h["nonleader_evals"], h["leaders"],
                h["orphans"], h["far_levels"]) == (
            o["filter_evals"], o["conflict_pairs"], o["nonleader_evals"], o["leaders"],
            o["orphans_dropped"], o["far_levels"])


@pytest.mark.parametrize("graphs", ["0", "1"])
def test_build_with_and_without_recorded_batches(hip, monkeypatch, graphs):
    # full batches replay one recorded launch graph (or are launched one by
    # one): the same tree and the same per-batch counters either way
    monkeypatch.setenv("COVERTREEX_HIP_BUILD_GRAPHS", graphs)
    case = dict(n=5000, d=3, m=64, kt=1, chunk=256, gs=1.0)
    x, v, p, ot, ht = _tree_pair(hip, case)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    tel = ht.build_telemetry()
    assert [(b["start"], b["size"]) for b in tel] == [(o["start"], o["size"])
                                                      for o in ot.batch_stats]
    assert [b["scan_evals"] for b in tel] == [o["filter_evals"] for o in ot.batch_stats]
    assert [b["leaders"] for b in tel] == [o["leaders"] for o in ot.batch_stats]


@pytest.mark.parametrize("max_work", ["0", "3000", "1000000000000000"])
@pytest.mark.parametrize("case", [BUILD_CASES[0], BUILD_CASES[6],
                                  dict(n=6000, d=3, m=32, kt=0, chunk=128, gs=1.0)])
def test_build_fused_and_wide_segments(hip, monkeypatch, case, max_work):
    # batches run in segments whose mode (fused or wide levels) follows the
    # work measured so far; the threshold moves the switches (0: wide after
    # the first segment that did any fused-level work; huge: always fused),
    # never the tree or the per-batch counters
    monkeypatch.setenv("COVERTREEX_HIP_FUSED_MAX_WORK", max_work)
    print(f"case {case} max work {max_work}", flush=True)
    x, v, p, ot, ht = _tree_pair(hip, case)
    assert np.array_equal(ht.get_parents(), ot.parents)
    assert np.array_equal(ht.get_children(), ot.children)
    assert np.array_equal(ht.get_next_node(), ot.next)
    assert np.array_equal(ht.get_levels(), ot.levels)
    tel = ht.build_telemetry()
    assert [(b["scan_evals"], b["mis_evals"], b["nonleader_evals"], b["leaders"])
            for b in tel] == [(o["filter_evals"], o["conflict_pairs"], o["nonleader_evals"],
                               o["leaders"]) for o in ot.batch_stats]
