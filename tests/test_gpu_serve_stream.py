"""The serving-stream workload at its stated sizes: requests drawn from the
512-request stream (N log-uniform in [2,048, 131,072], D in {2,3,5},
Matern/RBF with lengthscale U[0.5, 2], m in {64,128,256}, k in {10,30,50},
predecessor mode), including smooth-field requests (RBF, D=2, lengthscale
> 1: p near its floor on most rows) and requests above 100K rows.

Each request's result is the same alone and in the mix (bit for bit), and a
sample of each request's rows agrees with the FP64 brute force on the same
stored inputs (sets equal except ties within 1e-6, distance error <= 1e-5,
the stated bound covering the error)."""

import numpy as np
import pytest

from test_gpu_mixed import _check

pytestmark = pytest.mark.gpu

# the first 32 requests of the stream (seven smooth fields, N up to 97,602),
# request 39 (N=103,754, D=5) and request 351 (smooth field, N=115,540)
PICK = list(range(32)) + [39, 351]


def test_stream_requests_alone_in_mix_and_vs_fp64(hip):
    from covertreex_amd.residual import build_residual_backend
    from covertreex_amd.serve import Server, make_stream, run_poisson, run_request

    stream = make_stream(max(PICK) + 1, seed=0)
    reqs = [stream[i] for i in PICK]
    del stream
    assert any(r.n > 100_000 for r in reqs)
    assert any(r.kernel_type == 0 and r.coords.shape[1] == 2 and r.lengthscale > 1.0
               and r.n > 100_000 for r in reqs)
    alone = []
    for i, r in zip(PICK, reqs):
        print(f"alone: request {i} n={r.n} d={r.coords.shape[1]} kt={r.kernel_type} "
              f"ls={r.lengthscale:.2f} m={r.m} k={r.k}", flush=True)
        alone.append(run_request(r)[:3])
    server = Server(max_inflight=4, hbm_budget_bytes=64 << 30)
    try:
        print("mix: Poisson arrivals at 40 requests/s", flush=True)
        mixed = run_poisson(server, reqs, rate_per_s=40.0, seed=2)
    finally:
        server.close()
    assert server.peak_inflight >= 2
    rng = np.random.default_rng(7)
    for i, r, a, mres in zip(PICK, reqs, alone, mixed):
        assert np.array_equal(a[0], mres.indices), f"request {i}: indices differ in the mix"
        assert np.array_equal(a[1].view(np.uint32), mres.distances.view(np.uint32)), (
            f"request {i}: distances differ in the mix")
        assert np.array_equal(a[2].view(np.uint32), mres.errors.view(np.uint32))
        valid = mres.indices >= 0
        assert np.array_equal(valid.sum(1), np.minimum(r.k, np.arange(r.n)))
        # the stored inputs of the request (the device Nystrom is deterministic)
        b = build_residual_backend(r.coords, seed=r.seed, inducing_count=r.m,
                                   variance=r.variance, lengthscale=r.lengthscale,
                                   kernel_type=r.kernel_type)
        q = np.unique(np.concatenate([[0, 1, 2, r.n - 1],
                                      rng.choice(r.n, 28, replace=False)]))
        print(f"fp64: request {i} rows {q.size}", flush=True)
        worst = _check(mres.indices[q], mres.distances[q], mres.errors[q], r.coords,
                       b.v_matrix, b.p_diag, q, r.k, True, r.kernel_type, var=r.variance,
                       ls=float(np.float32(r.lengthscale)))
        assert worst <= 1e-5
