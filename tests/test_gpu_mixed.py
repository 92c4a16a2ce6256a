"""Mixed-precision engine vs the FP64 brute-force oracle on the same stored
inputs: neighbour sets equal except at distance ties within 1e-6, distances
within 1e-5 of FP64, and every reported error bound covers the actual error."""

import numpy as np
import pytest

import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu

TIE = 1e-6
DIST_TOL = 1e-5


def _check(idx, dist, err, x, v, p, q, k, pred, kt, var=1.0, ls=1.0):
    bi, bd, bn = fp64.brute_knn_fp64(x, v, p, q, k, predecessor=pred, kernel_type=kt,
                                     variance=var, lengthscale=ls)
    worst = 0.0
    for r in range(q.size):
        got = idx[r][idx[r] >= 0]
        ref = bi[r][bi[r] >= 0]
        assert got.size == ref.size, f"row {r}: {got.size} vs {ref.size} neighbours"
        if got.size == 0:
            continue
        d64 = fp64.residual_dist_fp64(x, v, p, [q[r]], got, kernel_type=kt, variance=var,
                                      lengthscale=ls)[0]
        e = np.abs(dist[r][: got.size].astype(np.float64) - d64)
        worst = max(worst, float(e.max()))
        assert np.all(e <= DIST_TOL), f"row {r}: distance error {e.max()}"
        assert np.all(e <= err[r][: got.size].astype(np.float64) + 2e-7), f"row {r}: bound"
        if set(got.tolist()) != set(ref.tolist()):
            kth = bd[r][ref.size - 1]
            diff = set(got.tolist()) ^ set(ref.tolist())
            dd = fp64.residual_dist_fp64(x, v, p, [q[r]], np.array(sorted(diff)),
                                         kernel_type=kt, variance=var, lengthscale=ls)[0]
            assert np.all(np.abs(dd - kth) <= TIE), f"row {r}: sets differ beyond ties"
    return worst


@pytest.mark.parametrize("n,d,m,kt,k,pred,dist", [
    (4096, 3, 64, 1, 16, False, "normal"),   # C1
    (4096, 3, 64, 1, 16, True, "normal"),
    (3000, 2, 128, 0, 10, True, "normal"),
    (2500, 5, 96, 0, 30, True, "uniform"),
    (700, 3, 32, 1, 50, True, "normal"),     # k near the early rows' counts
    (300, 3, 300, 1, 8, False, "normal"),    # m = N: every row inducing
])
def test_mixed_exact_vs_fp64(hip, n, d, m, kt, k, pred, dist):
    from covertreex_amd.native import HipTree

    print(f"n={n} d={d} m={m} kt={kt} k={k} pred={pred}", flush=True)
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((n, d)) if dist == "normal" else rng.uniform(-3, 3, (n, d)))
    x = x.astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=kt)
    q = np.arange(n) if n <= 1000 else rng.choice(n, 600, replace=False)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=pred, mode="mixed", return_error=True)
    worst = _check(idx, dst, err, x, v, p, q, k, pred, kt)
    print(f"max |d - d64| = {worst:.3e}", flush=True)
    if pred:
        for r, qi in enumerate(q):
            assert np.all(idx[r][idx[r] >= 0] < qi)
            assert (idx[r] >= 0).sum() == min(k, qi)


def test_mixed_ard_lengthscale(hip):
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(5)
    n, d = 2000, 5
    x = rng.uniform(-3, 3, (n, d)).astype(np.float32)
    ls = np.exp(rng.uniform(np.log(0.5), np.log(2.0), d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=128, seed=0, kernel_type=0, lengthscale=ls)
    t = HipTree.build_residual(v, p, x, 1.0, ls, kernel_type=0)
    q = rng.choice(n, 300, replace=False)
    idx, dst, err = t.knn_residual(q, 30, predecessor_mode=True, mode="mixed", return_error=True)
    _check(idx, dst, err, x, v, p, q, 30, True, 0, ls=ls.astype(np.float64))


def test_mixed_query_order_and_duplicates(hip):
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(2)
    x = rng.standard_normal((1500, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=64, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    q = np.array([1499, 3, 3, 0, 700, 1, 1200])
    a_i, a_d = t.knn_residual(q, 12, predecessor_mode=True, mode="mixed")
    for r, qi in enumerate(q):
        b_i, b_d = t.knn_residual(np.array([qi]), 12, predecessor_mode=True, mode="mixed")
        assert np.array_equal(a_i[r], b_i[0]) and np.array_equal(a_d[r], b_d[0])
    assert np.all(a_i[3] == -1)
    assert set(a_i[5][a_i[5] >= 0]) == {0}


def test_c2_shape_sample(hip):
    # N=32768, D=3, Matern l=1, m=256, k=50, predecessor mode
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(0)
    x = rng.standard_normal((32768, 3)).astype(np.float32)
    v, p, _, _, _, ind = fp64.nystrom(x, inducing_count=256, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    q = np.arange(32768)
    t.set_query_telemetry(True)
    idx, dst, err = t.knn_residual(q, 50, predecessor_mode=True, mode="mixed", return_error=True)
    # the inducing rows (p at the jitter floor) overflow the screen's lists and
    # take the exact-product screen: a quarter of them join the random sample
    tel = t.last_query_telemetry()
    assert tel["rescreened_queries"] > 100
    assert tel["hard_queries"] < tel["rescreened_queries"] // 4
    sample = np.concatenate([np.random.default_rng(1).choice(32768, 256, replace=False),
                             ind[1::4]])
    worst = _check(idx[sample], dst[sample], err[sample], x, v, p, sample, 50, True, 1)
    print(f"C2 sample max |d - d64| = {worst:.3e}", flush=True)
    # size-independent properties over all rows
    valid = idx >= 0
    assert np.array_equal(valid.sum(1), np.minimum(50, q))
    assert np.all(np.where(valid, idx, -1) < q[:, None])
    dd = np.where(valid, dst, 2.0)
    assert np.all(np.diff(dd, axis=1) >= 0)
    # the certified error bound of every returned distance, all 32,768 rows
    eb = float(err[valid].max())
    print(f"C2 max error bound over all rows = {eb:.3e}", flush=True)
    assert eb <= DIST_TOL


def test_smooth_field_rescreened(hip):
    # RBF, l = 1.5 in D = 2 with m = 256: p ~ 1e-7 on nearly every row, so
    # nearly every query overflows the BF16 screen and is listed again by the
    # exact-product screen (at larger N some lists still overflow, with many
    # candidates tied at distance 0, and take the exhaustive path)
    from covertreex_amd.native import HipTree

    n, k = 8000, 30
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, 2)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=256, seed=0, kernel_type=0, lengthscale=1.5)
    t = HipTree.build_residual(v, p, x, 1.0, np.full(1, 1.5, np.float32), kernel_type=0)
    t.set_query_telemetry(True)
    q = np.arange(n)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=True, mode="mixed", return_error=True)
    tel = t.last_query_telemetry()
    print(f"smooth field: {tel['rescreened_queries']} rescreened, "
          f"{tel['hard_queries']} scored exhaustively", flush=True)
    assert tel["rescreened_queries"] > n // 2
    sample = rng.choice(n, 200, replace=False)
    worst = _check(idx[sample], dst[sample], err[sample], x, v, p, sample, k, True, 0, ls=1.5)
    print(f"smooth field max |d - d64| = {worst:.3e}", flush=True)
    valid = idx >= 0
    assert np.array_equal(valid.sum(1), np.minimum(k, q))


def test_tied_rows_exhaustive_path(hip):
    # 10,000 copies of one point: a query among them has as many predecessors
    # at one and the same distance, which no certified interval separates.
    # Between 2,048 and 8,192 of them the exact-product screen lists the
    # query again with 8,192 slots; beyond that it is scored against all its
    # candidates
    from covertreex_amd.native import HipTree

    n, k = 24000, 20
    rng = np.random.default_rng(8)
    x = rng.standard_normal((n, 3)).astype(np.float32)
    x[2000:12000] = x[1999]
    v, p, *_ = fp64.nystrom(x, inducing_count=64, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    t.set_query_telemetry(True)
    q = np.arange(n)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=True, mode="mixed", return_error=True)
    tel = t.last_query_telemetry()
    print(f"tied rows: {tel['rescreened_queries']} rescreened, "
          f"{tel['hard_queries']} scored exhaustively", flush=True)
    # queries 10,192 .. 11,999 have more than 8,192 tied predecessors
    assert 1700 <= tel["hard_queries"] <= 1900
    sample = np.concatenate([np.arange(11900, 12000), np.arange(7000, 7100),
                             rng.choice(n, 100, replace=False)])
    _check(idx[sample], dst[sample], err[sample], x, v, p, sample, k, True, 1)
    valid = idx >= 0
    assert np.array_equal(valid.sum(1), np.minimum(k, q))


@pytest.mark.parametrize("pred,k", [(False, 100), (True, 24)])
def test_overflowed_queries_rescreened(hip, pred, k):
    # C1 shape (N=4096, D=3, Matern, m=64): the inducing rows overflow the
    # BF16 screen and are listed again by the exact-product screen, then
    # rescored with the others (k = 100 takes the three-entries-per-lane
    # top-k; without the mask the query's own row is a candidate)
    from covertreex_amd.native import HipTree

    x = np.random.default_rng(0).standard_normal((4096, 3)).astype(np.float32)
    v, p, _, _, _, ind = fp64.nystrom(x, inducing_count=64, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    t.set_query_telemetry(True)
    q = np.arange(4096)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=pred, mode="mixed", return_error=True)
    tel = t.last_query_telemetry()
    assert tel["rescreened_queries"] > 0
    sample = np.concatenate([ind, np.random.default_rng(2).choice(4096, 64, replace=False)])
    worst = _check(idx[sample], dst[sample], err[sample], x, v, p, sample, k, pred, 1)
    print(f"rescreened {tel['rescreened_queries']} exhaustive {tel['hard_queries']} "
          f"max |d - d64| = {worst:.3e}", flush=True)


@pytest.mark.parametrize("n,d,m,kt,k,pred", [
    (3000, 4, 50, 1, 200, True),    # generic screen (d = 4), unstaged query row (m % 4 != 0),
                                    # seven kept entries per lane (k in 193..256)
    (2000, 8, 600, 0, 40, False),   # d = 8 and m > 512
])
def test_mixed_generic_paths_vs_fp64(hip, n, d, m, kt, k, pred):
    from covertreex_amd.native import HipTree

    print(f"n={n} d={d} m={m} kt={kt} k={k} pred={pred}", flush=True)
    rng = np.random.default_rng(17)
    x = rng.standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=kt)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=kt)
    q = rng.choice(n, 300, replace=False)
    idx, dst, err = t.knn_residual(q, k, predecessor_mode=pred, mode="mixed", return_error=True)
    _check(idx, dst, err, x, v, p, q, k, pred, kt)


def test_mixed_variance_mismatched_to_v(hip):
    # V and p built with variance 1, the tree with variance 2: |rho| exceeds 1
    # for many pairs and is clamped (the FP64 oracle clips it the same way);
    # the screen's bounds must stay inside their histogram
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(23)
    n, k = 3000, 20
    x = rng.standard_normal((n, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=64, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 2.0, np.ones(1, np.float32), kernel_type=1)
    q = rng.choice(n, 300, replace=False)
    for pred in (True, False):
        idx, dst, err = t.knn_residual(q, k, predecessor_mode=pred, mode="mixed",
                                       return_error=True)
        _check(idx, dst, err, x, v, p, q, k, pred, 1, var=2.0)


def test_mixed_floored_p_pair(hip):
    # Every p at 1e-9, so p_q p_j (9.9999994e-19 in FP32) sits below the
    # 1e-18 floor and the denominator is exactly 1e-9.  Coordinates coincide
    # (k = var) and var = 1e-8 with V rows of norm ~0.95e-4 in a narrow cone,
    # so the numerators var - V_q.V_j sit just below 1e-9 and 1 - |rho| spreads
    # over [0, 2e-4]: distances up to 0.011, where taking the FP32 literal
    # 1e-18f (1.0000000458e-18) for the floor would move them by up to 1.7e-6,
    # far beyond the reported bound the test holds them to.
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(29)
    # the library stores var in FP32; the FP64 oracle takes that stored value
    n, k, m, var = 1500, 10, 8, float(np.float32(1e-8))
    x = np.zeros((n, 3), np.float32)
    r = np.sqrt(0.9) * (1.0 + rng.uniform(0.0, 1e-5, n))
    th = rng.uniform(0.0, 1e-3, n)
    v = np.zeros((n, m), np.float32)
    v[:, 0] = 1e-4 * r * np.cos(th)
    v[:, 1] = 1e-4 * r * np.sin(th)
    p = np.full(n, 1e-9, np.float32)
    t = HipTree.build_residual(v, p, x, var, np.ones(1, np.float32), kernel_type=1)
    q = rng.choice(n, 200, replace=False)
    for pred in (True, False):
        idx, dst, err = t.knn_residual(q, k, predecessor_mode=pred, mode="mixed",
                                       return_error=True)
        _check(idx, dst, err, x, v, p, q, k, pred, 1, var=var)
        got = dst[idx >= 0]
        assert got.max() > 5e-3 and np.any((got > 0) & (got < 2e-3))


def test_mixed_out_of_range_queries(hip):
    # out-of-range query indices read tree node 0 (src/lib.rs:414-418) and,
    # as usize bounds, admit every row in predecessor mode
    from covertreex_amd.native import HipTree

    rng = np.random.default_rng(31)
    n, k = 2000, 12
    x = rng.standard_normal((n, 3)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=64, seed=0, kernel_type=1)
    t = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    node0 = int(t.get_order()[0])
    want_i, want_d = t.knn_residual(np.array([node0]), k, predecessor_mode=False, mode="mixed")
    for pred in (True, False):
        gi, gd = t.knn_residual(np.array([-1, n + 5, node0]), k, predecessor_mode=pred,
                                mode="mixed")
        for r in (0, 1):
            assert np.array_equal(gi[r], want_i[0]) and np.array_equal(gd[r], want_d[0])


@pytest.mark.parametrize("pred", [True, False])
def test_screen_segment_lengths_same_results(hip, pred):
    # the screen's segment length (its tile shape, set by the tuner) changes
    # which workgroup scores a pair, never the neighbours, distances or bounds
    from covertreex_amd.native import HipTree

    n, d, m, k = 5000, 3, 128, 20
    x = np.random.default_rng(4).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=1)
    tree = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    q = np.arange(n)
    outs = []
    for seg in (0, 256, 768, 4096, 16384):
        tree.set_screen_segment(seg)
        outs.append(tree.knn_residual(q, k, predecessor_mode=pred, mode="mixed",
                                      return_error=True))
        # default for mp = 128: 4096, halved while the grid holds fewer than
        # 512 workgroups (79 query tiles: down to 512)
        assert tree.screen_segment() == (seg or 512)
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1].view(np.uint32), outs[0][1].view(np.uint32))
        assert np.array_equal(o[2].view(np.uint32), outs[0][2].view(np.uint32))
    qs = q[::37]
    _check(outs[0][0][qs], outs[0][1][qs], outs[0][2][qs], x, v, p, qs, k, pred, 1)
    with pytest.raises(ValueError):
        tree.set_screen_segment(300)


@pytest.mark.parametrize("pred", [True, False])
def test_launch_groups_same_results(hip, monkeypatch, pred):
    # a small code budget splits the screen into launch groups (each
    # group's screen, tau and filter over its query tiles, one code buffer
    # reused); the results are those of one group
    from covertreex_amd.native import HipTree

    n, d, m, k = 5000, 3, 128, 20
    x = np.random.default_rng(5).standard_normal((n, d)).astype(np.float32)
    v, p, *_ = fp64.nystrom(x, inducing_count=m, seed=0, kernel_type=1)
    tree = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    q = np.arange(n)
    one = tree.knn_residual(q, k, predecessor_mode=pred, mode="mixed", return_error=True)
    # 79 query tiles of 64 queries; a tile's codes take n_seg * seg * 64 bytes
    tile_bytes = ((n + 511) // 512) * 512 * 64
    for tiles in (1, 2, 5, 40):
        monkeypatch.setenv("COVERTREEX_HIP_CODE_BUDGET", str(tiles * tile_bytes))
        got = tree.knn_residual(q, k, predecessor_mode=pred, mode="mixed", return_error=True)
        assert np.array_equal(got[0], one[0])
        assert np.array_equal(got[1].view(np.uint32), one[1].view(np.uint32))
        assert np.array_equal(got[2].view(np.uint32), one[2].view(np.uint32))
    qs = q[::41]
    _check(one[0][qs], one[1][qs], one[2][qs], x, v, p, qs, k, pred, 1)
