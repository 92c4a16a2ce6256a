#!/usr/bin/env python3
"""Benchmark: residual kNN queries/s at N=32768, D=3, k=50, predecessor mode.

Workload (one step): the Vecchia-shape residual kNN over every point of a
32,768-point, 3-D N(0,1) dataset with a Matern-5/2 (lengthscale 1) Nystrom
factor of rank 256, k=50, neighbours j < i, answered by the mixed-precision
engine from inputs already resident in HBM.  Beside it, in the same run: the
plain FP32 GPU form (the reference's search semantics and f32 arithmetic, one
wavefront per query) and the C oracle's restatement of the reference search on
the host cores over a bounded query sample.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.

`--workload c1|c3|c4` measures one of the other workloads instead (its own
JSON line, no CPU leg), for profiling: c1 the 4,096-point, rank-64, k=16 query
of every point (not predecessor); c3 the Euclidean build of 1,048,576 points
in D=8 from a 64-component Gaussian mixture and the exact kNN (k=32) of 65,536
separate queries; c4 the 262,144-point, D=5, anisotropic-RBF, rank-512, k=30
predecessor query of every point (with its Nystrom factor and build timed).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, D, M, K = 32768, 3, 256, 50
MFMA_BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
# HBM bytes per screen_kernel<3, 1> launch on this workload: rocprofv3
# FETCH_SIZE (x2.000, calibrated on a 1 GiB copy) + WRITE_SIZE (x1.000),
# recorded in profiles/mixed_knn_c2_n32768_m256_k50.md
SCREEN_PASS1_HBM_BYTES = 2.894e9


def make_inputs():
    from covertreex_amd.nystrom import nystrom_device
    from covertreex_amd.residual import inducing_indices

    x = np.random.default_rng(0).standard_normal((N, D)).astype(np.float32)
    ind = inducing_indices(N, M, 0)
    v, p = nystrom_device(x.astype(np.float64), ind, variance=1.0, lengthscale=1.0,
                          kernel_type=1)
    return x, v, p


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x, v, p, sample_q: int):
    """The oracle's restatement of the reference search (exact semantics) on
    the host cores over a bounded sample of queries; build excluded.  The
    timed copy is compiled here for this host (-O3 -march=native, the same
    floating-point flags); the in-tree build is used if that fails."""
    import tempfile

    import oracle
    import oracle.lib as ol

    native = os.path.join(tempfile.mkdtemp(prefix="cx_oracle_"), "liboracle_native.so")
    built = ol.build_native(native)
    if built:
        ol.use_library(native)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t = oracle.build_residual_tree(x, v, p, kernel_type=1)
    q = np.random.default_rng(3).choice(N, sample_q, replace=False)
    t0 = time.perf_counter()
    _, _, evals = oracle.knn_residual(t, q, K, predecessor=True, budget=(), threads=threads)
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.knn_residual(t, q, K, predecessor=True, budget="default", threads=threads)
    el_b = time.perf_counter() - t0
    q1 = q[: max(1, sample_q // 8)]
    t0 = time.perf_counter()
    oracle.knn_residual(t, q1, K, predecessor=True, budget=(), threads=1)
    el_1 = time.perf_counter() - t0
    return {"value": round(sample_q / el, 1), "unit": "queries/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample_q} random queries of the same workload, exact search "
                      f"(empty budget schedule), {threads} OpenMP threads",
            "cpu_model": _cpu_model(),
            "compiler_flags": ("-O3 -march=native -fopenmp -ffp-contract=off" if built
                               else "-O3 -fopenmp -ffp-contract=off (in-tree build)"),
            "one_thread_qps": round(q1.size / el_1, 1),
            "one_thread_sample": f"{q1.size} of the same queries, exact search, 1 thread",
            "budgeted_default_qps": round(sample_q / el_b, 1),
            "mean_evals_per_query": float(evals.mean())}


def measure_peaks(torch, dev) -> dict:
    """What this device sustains on two plain library operations: a 1 GiB
    device copy (read + write bytes / time) and a dense 8192^3 BF16 GEMM
    (hipBLASLt through torch.mm), each timed with events over repeated calls."""
    out = {}
    a = torch.ones((1 << 28,), dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    out["hbm_copy_GBps"] = round(5 * 2 * a.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    del a, b
    n = 8192
    x = torch.randn((n, n), dtype=torch.bfloat16, device=dev)
    y = torch.randn((n, n), dtype=torch.bfloat16, device=dev)
    z = torch.mm(x, y)
    e0.record()
    for _ in range(10):
        z = torch.mm(x, y)
    e1.record()
    torch.cuda.synchronize()
    out["bf16_gemm_TFLOPs"] = round(10 * 2 * n ** 3 / (e0.elapsed_time(e1) * 1e-3) / 1e12, 1)
    del x, y, z
    torch.cuda.empty_cache()
    return out


def _timed(fn, steps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def other_workload(name: str, steps: int, warmup: int) -> dict:
    """The c1, c3 and c4 lines (see the module docstring)."""
    import torch

    from covertreex_amd.native import HipTree

    dev = torch.device("cuda:0")
    if name == "c3":
        from covertreex_amd.euclidean import EuclideanHipTree

        centres = np.random.default_rng(0).random((64, 8))

        def mixture(n, seed):
            rng = np.random.default_rng(seed)
            lab = rng.integers(0, 64, n)
            return (centres[lab] + 0.02 * rng.standard_normal((n, 8))).astype(np.float32)

        pts, qs = mixture(1 << 20, 0), mixture(65536, 1)
        t0 = time.perf_counter()
        tree = EuclideanHipTree.build(pts, chunk_size=512)
        build_s = time.perf_counter() - t0
        for _ in range(warmup):
            tree.knn_points(qs, 32, mode="exact")
        el = _timed(lambda: tree.knn_points(qs, 32, mode="exact"), steps, torch)
        return {"metric": "euclidean exact kNN queries/s @N=1M,D=8,k=32 (host buffers)",
                "value": round(qs.shape[0] / el, 1), "unit": "queries/s",
                "ms_per_step": round(el * 1e3, 3), "steps": steps, "warmup": warmup,
                "build_s": round(build_s, 3), "dtype": "f32",
                "config": {"workload": "euclidean GMM 64 components sigma 0.02: N=1048576 "
                                       "D=8, 65536 separate queries, k=32, batches of 512"}}
    from covertreex_amd.residual import build_residual_backend

    if name == "c1":
        n, d, m, k, pred, kt = 4096, 3, 64, 16, False, 1
        x = np.random.default_rng(0).standard_normal((n, d)).astype(np.float32)
        ls = np.ones(1, np.float32)
    else:
        n, d, m, k, pred, kt = 262144, 5, 512, 30, True, 0
        x = np.random.default_rng(0).uniform(-3, 3, (n, d)).astype(np.float32)
        ls = np.exp(np.random.default_rng(0).uniform(np.log(0.5), np.log(2.0), d))
    t0 = time.perf_counter()
    b = build_residual_backend(x, seed=0, inducing_count=m, variance=1.0, lengthscale=ls,
                               kernel_type=kt)
    nystrom_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tree = HipTree.build_residual(b.v_matrix, b.p_diag, x, 1.0, np.asarray(ls, np.float32),
                                  kernel_type=kt)
    build_s = time.perf_counter() - t0
    q = torch.arange(n, dtype=torch.int64, device=dev)
    o_idx = torch.empty((n, k), dtype=torch.int64, device=dev)
    o_dst = torch.empty((n, k), dtype=torch.float32, device=dev)
    o_err = torch.empty((n, k), dtype=torch.float32, device=dev)

    def step():
        tree.knn_residual_device(q.data_ptr(), n, k, o_idx.data_ptr(), o_dst.data_ptr(),
                                 o_err.data_ptr(), predecessor_mode=pred, mode="mixed")

    for _ in range(warmup):
        step()
    el = _timed(step, steps, torch)
    valid = o_idx >= 0
    tree.set_profiling(True)
    step()
    phases = dict(zip(HipTree.PHASES["mixed"], tree.last_profile()))
    tree.set_profiling(False)
    label = ("residual kNN queries/s @N=4096,D=3,k=16" if name == "c1"
             else "residual kNN queries/s @N=262144,D=5,k=30 predecessor")
    return {"metric": label, "value": round(n / el, 1), "unit": "queries/s",
            "ms_per_step": round(el * 1e3, 3), "steps": steps, "warmup": warmup,
            "nystrom_s": round(nystrom_s, 3), "build_s": round(build_s, 3),
            "max_err_bound": float(o_err[valid].max().item()),
            "phases_ms": {k_: round(v_, 4) for k_, v_ in phases.items()},
            "dtype": "bf16-hilo-mfma/f32-compensated",
            "config": {"workload": (f"residual N={n} D={d} m={m} k={k} "
                                    f"{'predecessor' if pred else 'all neighbours'} "
                                    f"{'Matern52 l=1' if kt == 1 else 'ARD RBF'}")}}


def build_timing(reps: int) -> dict:
    """Wall time of the C2 build call (host V, p and coords in; the tree
    resident in HBM out), one untimed build first."""
    from covertreex_amd.native import HipTree

    x, v, p = make_inputs()
    ms = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        tree = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
        if r:
            ms.append(round((time.perf_counter() - t0) * 1e3, 3))
        tree.close()
    return {"build_ms": ms, "median_ms": float(np.median(ms)),
            "config": {"workload": f"C2 build: N={N} D={D} m={M}, chunk 512, Morton order"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=15.0)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--workload", choices=("c2", "c1", "c3", "c4"), default="c2",
                    help="c2 (default) is the headline line; c1/c3/c4 print their own")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--build-reps", type=int, default=0,
                    help="only time R builds of the C2 tree from host V, p and coords "
                         "(after one untimed build)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration:
        # a plain device copy of known size: what FETCH_SIZE / WRITE_SIZE read
        # for it scales those counters to bytes
        nbytes = 1 << 30
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))
        return 0

    if args.build_reps > 0:
        print(json.dumps(build_timing(args.build_reps)))
        return 0

    if args.workload != "c2":
        print(json.dumps(other_workload(args.workload, args.steps, args.warmup)))
        return 0

    from covertreex_amd.native import HipTree

    dev = torch.device("cuda:0")
    peaks = measure_peaks(torch, dev)
    x, v, p = make_inputs()
    t0 = time.perf_counter()
    tree = HipTree.build_residual(v, p, x, 1.0, np.ones(1, np.float32), kernel_type=1)
    build_s = time.perf_counter() - t0

    q = torch.arange(N, dtype=torch.int64, device=dev)
    o_idx = torch.empty((N, K), dtype=torch.int64, device=dev)
    o_dst = torch.empty((N, K), dtype=torch.float32, device=dev)
    o_err = torch.empty((N, K), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step(mode="mixed"):
        tree.knn_residual_device(q.data_ptr(), N, K, o_idx.data_ptr(), o_dst.data_ptr(),
                                 o_err.data_ptr(), predecessor_mode=True, mode=mode)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    qps = N * args.steps / elapsed

    # largest certified error bound over every returned distance of the last
    # timed call (the bound holds against the FP64 evaluation on the same inputs)
    valid = o_idx >= 0
    max_err_bound = float(o_err[valid].max().item()) if bool(valid.any()) else 0.0

    # per-phase device timing on the tree's own stream (HIP events)
    tree.set_profiling(True)
    prof = []
    for _ in range(3):
        step()
        prof.append(tree.last_profile())
    tree.set_profiling(False)
    prof = np.array(prof)
    phases = dict(zip(HipTree.PHASES["mixed"], np.median(prof, axis=0).tolist()))

    # plain FP32 GPU form (the denominator of the speedup)
    step("reference")
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(args.plain_steps):
        step("reference")
    torch.cuda.synchronize()
    plain_qps = N * args.plain_steps / (time.perf_counter() - s0)

    serve = None
    if args.serve_requests > 0:
        from covertreex_amd.serve import Server, latency_summary, make_stream, run_poisson

        print(f"bench.py: serving stream, {args.serve_requests} requests at "
              f"{args.serve_rate}/s", file=sys.stderr, flush=True)
        reqs = make_stream(args.serve_requests, seed=0, n_min=2048, n_max=131072)
        server = Server(max_inflight=args.serve_inflight, hbm_budget_bytes=64 << 30)
        try:
            res = run_poisson(server, reqs, rate_per_s=args.serve_rate, seed=0)
        finally:
            server.close()
        serve = {k_: (round(v_, 3) if isinstance(v_, float) else v_)
                 for k_, v_ in latency_summary(res).items()}
        serve.update({"rate_per_s": args.serve_rate, "max_inflight": args.serve_inflight,
                      "peak_inflight": server.peak_inflight,
                      "workload": "Poisson arrivals; N log-uniform [2048, 131072], D in {2,3,5}, "
                                  "Matern/RBF l~U[0.5,2], m in {64,128,256}, k in {10,30,50}, "
                                  "predecessor; build + all-points kNN per request"})

    cpu = None
    evals = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, v, p, args.cpu_sample)
        evals = cpu.get("mean_evals_per_query")
    # algorithmic work per query row: E_q (x) (2m + 2D + 20) flops with E_q the
    # reference exact search's distance evaluations on this tree
    e_q = evals if evals else float(N)
    flops_per_q = e_q * (2 * M + 2 * D + 20)
    screen_ms = phases["screen_pass1"]
    achieved = N * flops_per_q / (screen_ms * 1e-3) / 1e12
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "residual kNN queries/s @N=32k,D=3,k=50 vs plain GPU; p50/p99 serve latency",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-hilo-mfma/f32-compensated",
        "data": "synthetic",
        "config": {"workload": "vecchia residual kNN: N=32768 D=3 N(0,1) Matern52 l=1 "
                               "Nystrom m=256 k=50 predecessor all queries",
                   "n": N, "d": D, "m": M, "k": K, "predecessor": True},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "max_err_bound": max_err_bound,
        "plain_f32_gpu_qps": round(plain_qps, 1),
        "speedup_vs_plain_f32_gpu": round(qps / plain_qps, 2),
        "build_s": round(build_s, 3),
        "phases_ms": {k_: round(v_, 4) for k_, v_ in phases.items()},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2),
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4),
                     "traffic": SCREEN_PASS1_HBM_BYTES,
                     # the launch's calibrated HBM bytes over its measured time
                     "hbm_frac": round(SCREEN_PASS1_HBM_BYTES / (screen_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "traffic_source": "profiles/mixed_knn_c2_n32768_m256_k50.md",
                     "kernel": "screen_kernel<3, 1>",
                     "algorithmic_flops_per_launch": N * flops_per_q},
        "cpu_baseline": cpu,
        "serve": serve,
        "measured_peaks": dict(peaks, screen_frac_of_measured_bf16_gemm=round(
            achieved / peaks["bf16_gemm_TFLOPs"], 4) if peaks.get("bf16_gemm_TFLOPs") else None),
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
