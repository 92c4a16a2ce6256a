"""Serving loop: requests overlap on the device and each result equals the
same request run alone."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_results_identical_alone_and_in_mix(hip):
    from covertreex_amd.serve import Server, make_stream, run_poisson, run_request

    reqs = make_stream(12, seed=3, n_min=1024, n_max=8192)
    alone = [run_request(r)[:3] for r in reqs]
    server = Server(max_inflight=4, hbm_budget_bytes=8 << 30)
    try:
        mixed = run_poisson(server, reqs, rate_per_s=200.0, seed=1)
    finally:
        server.close()
    assert server.peak_inflight >= 2  # requests really overlapped
    for a, r in zip(alone, mixed):
        assert np.array_equal(a[0], r.indices)
        assert np.array_equal(a[1].view(np.uint32), r.distances.view(np.uint32))
    for req, r in zip(reqs, mixed):
        valid = r.indices >= 0
        assert np.array_equal(valid.sum(1), np.minimum(req.k, np.arange(req.n)))


def test_admission_by_footprint(hip):
    from covertreex_amd.serve import Request, Server

    x = np.random.default_rng(0).standard_normal((4096, 3)).astype(np.float32)
    req = Request(coords=x, k=10, m=64)
    need = req.footprint_bytes()
    # budget for one request at a time: the second waits for the first
    server = Server(max_inflight=4, hbm_budget_bytes=int(need * 1.5))
    try:
        f1 = server.submit(req)
        f2 = server.submit(Request(coords=x, k=10, m=64))
        r1, r2 = f1.result(), f2.result()
    finally:
        server.close()
    assert server.peak_inflight == 1
    assert r2.admitted >= r1.done - 1e-3
    assert np.array_equal(r1.indices, r2.indices)
    with pytest.raises(ValueError):
        Server(max_inflight=1, hbm_budget_bytes=10).submit(req)


def test_api_and_factory_surface(hip):
    # cover_tree(kernel=...).knn and CoverTree(Runtime).fit/knn: the reference's
    # predecessor invariants (tests/test_predecessor_mode.py:638-686)
    import covertreex_amd as cx
    from covertreex_amd.api import CoverTree, Residual, Runtime

    pts = np.random.default_rng(42).standard_normal((500, 3)).astype(np.float32)
    tree = cx.cover_tree(pts, kernel=cx.Matern52(lengthscale=1.0), engine="hip")
    nb, dist = tree.knn(k=30, predecessor_mode=True, return_distances=True)
    for i in range(500):
        valid = nb[i][nb[i] >= 0]
        assert np.all(valid < i) and valid.size == min(30, i)
    assert np.all(nb[0] == -1)
    assert set(nb[1][nb[1] >= 0]) == {0}
    rt = Runtime(metric="residual_correlation",
                 residual=Residual(kernel_type=1, inducing_count=128, seed=42))
    api = CoverTree(rt).fit(pts)
    nb2 = api.knn(np.arange(500).reshape(-1, 1), k=30, predecessor_mode=True)
    v128 = cx.cover_tree(pts, kernel=cx.Matern52(1.0), inducing_count=128).knn(
        k=30, predecessor_mode=True)
    assert np.array_equal(nb2, v128)
    ref = cx.cover_tree(pts, kernel=cx.Matern52(1.0), precision="reference-f32")
    nr = ref.knn(k=30, predecessor_mode=True)
    for i in range(500):
        assert np.all(nr[i][nr[i] >= 0] < i)
