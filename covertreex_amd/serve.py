"""Continuous-batching request loop for residual kNN (build + full kNN per request).

Each request brings a dataset (coords), a kernel (type, lengthscale,
variance), a Nystrom rank m, k and the Vecchia flag.  A request is admitted
when a worker slot is free and its HBM footprint fits the budget next to the
requests already on the device; otherwise it waits (FIFO).  Admitted requests
run concurrently: every tree owns its own HIP stream, the C-ABI calls release
the GIL, so builds and queries of different requests overlap on the GPU.
Trees stay resident until their request completes (or, with ``keep_trees``,
until evicted by the footprint budget).

The reference has no counterpart; its per-request semantics are those of
``cover_tree(coords, kernel=...).knn(k, predecessor_mode=True)``.
"""

from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .kernels import Matern52, RBF
from .native import HipTree
from .residual import build_residual_backend


@dataclass
class Request:
    coords: np.ndarray
    k: int
    m: int
    kernel_type: int = 1
    lengthscale: float = 1.0
    variance: float = 1.0
    predecessor: bool = True
    seed: int = 0
    precision: str = "mixed"
    tag: object = None
    arrival: float = 0.0  # seconds on the server clock; filled by submit()

    @property
    def n(self) -> int:
        return int(self.coords.shape[0])

    def footprint_bytes(self) -> int:
        """HBM held while the request runs: V (f32, tree order) + BF16 hi/lo
        copy + coords/p/tree arrays + per-query lists and outputs, plus upper
        bounds of the mixed kNN's work buffers: one-byte screen codes (at
        most 4 GiB), the exact-product screen's int8 slices and stored bounds
        (512 MiB), second-listing lists (at most 1 GiB, 32 KiB per query) and
        exhaustive-path partial lists (512 MiB)."""
        n, m, d, k = self.n, self.m, self.coords.shape[1], self.k
        mp = (m + 31) // 32 * 32
        tree = n * (4 * m + 4 * mp + 12 * d + 120)
        query = n * (2048 * 4 + 64 * 4 + 60 + 12 * k)
        nystrom = n * (8 * m + 8 * d) + 8 * m * m
        work = (min(4 << 30, n * n) + n * 6 * mp + (512 << 20) + min(1 << 30, n * 32768)
                + (512 << 20))
        return int(tree + query + nystrom + work)


@dataclass
class Result:
    indices: np.ndarray
    distances: np.ndarray
    errors: np.ndarray
    arrival: float
    admitted: float
    done: float
    build_s: float
    knn_s: float
    tag: object = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def run_request(req: Request) -> tuple[np.ndarray, np.ndarray, np.ndarray, float, float]:
    """One request end to end on the GPU: Nystrom, build, all-points kNN."""
    kern = (Matern52 if req.kernel_type == 1 else RBF)(lengthscale=req.lengthscale,
                                                       variance=req.variance)
    t0 = time.perf_counter()
    b = build_residual_backend(req.coords, seed=req.seed, inducing_count=req.m,
                               variance=kern.variance, lengthscale=kern.lengthscale,
                               kernel_type=kern.kernel_type)
    tree = HipTree.build_residual(b.v_matrix, b.p_diag, b.kernel_points_f32, b.variance,
                                  np.atleast_1d(np.float32(req.lengthscale)),
                                  kernel_type=req.kernel_type)
    t1 = time.perf_counter()
    idx, dist, err = tree.knn_residual(np.arange(req.n), req.k,
                                       predecessor_mode=req.predecessor, mode=req.precision,
                                       return_error=True)
    t2 = time.perf_counter()
    tree.close()
    return idx, dist, err, t1 - t0, t2 - t1


@dataclass
class _Pending:
    req: Request
    fut: Future
    bytes: int


class Server:
    """Admission by slots and HBM footprint; FIFO wait queue."""

    def __init__(self, *, max_inflight: int = 4, hbm_budget_bytes: int = 64 << 30):
        self.max_inflight = int(max_inflight)
        self.budget = int(hbm_budget_bytes)
        self._lock = threading.Condition()
        self._queue: deque[_Pending] = deque()
        self._inflight = 0
        self._used = 0
        self._closed = False
        self._t0 = time.perf_counter()
        self._threads: list[threading.Thread] = []
        self.peak_inflight = 0
        for i in range(self.max_inflight):
            th = threading.Thread(target=self._worker, name=f"cx-serve-{i}", daemon=True)
            th.start()
            self._threads.append(th)

    def now(self) -> float:
        return time.perf_counter() - self._t0

    def submit(self, req: Request) -> Future:
        need = req.footprint_bytes()
        if need > self.budget:
            raise ValueError(f"request needs {need} bytes of HBM, budget is {self.budget}")
        fut: Future = Future()
        req.arrival = self.now()
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append(_Pending(req, fut, need))
            self._lock.notify_all()
        return fut

    def _take(self):
        with self._lock:
            while True:
                if self._queue:
                    head = self._queue[0]
                    if self._used + head.bytes <= self.budget:
                        self._queue.popleft()
                        self._used += head.bytes
                        self._inflight += 1
                        self.peak_inflight = max(self.peak_inflight, self._inflight)
                        return head
                elif self._closed:
                    return None
                self._lock.wait(0.05)

    def _worker(self):
        while True:
            item = self._take()
            if item is None:
                return
            admitted = self.now()
            try:
                idx, dist, err, bs, ks = run_request(item.req)
                item.fut.set_result(Result(idx, dist, err, item.req.arrival, admitted,
                                           self.now(), bs, ks, item.req.tag))
            except BaseException as exc:  # noqa: BLE001 - handed to the caller
                item.fut.set_exception(exc)
            finally:
                with self._lock:
                    self._inflight -= 1
                    self._used -= item.bytes
                    self._lock.notify_all()

    def close(self):
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for th in self._threads:
            th.join()


def make_stream(n_requests: int, *, seed: int = 0, n_min: int = 2048, n_max: int = 131072):
    """The serving-stream workload: N log-uniform in [n_min, n_max], D in
    {2,3,5}, Matern or RBF with lengthscale U[0.5,2], m in {64,128,256},
    k in {10,30,50}, predecessor mode."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_requests):
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        d = int(rng.choice([2, 3, 5]))
        kt = int(rng.integers(0, 2))
        ls = float(rng.uniform(0.5, 2.0))
        m = int(rng.choice([64, 128, 256]))
        k = int(rng.choice([10, 30, 50]))
        x = np.random.default_rng(seed * 100003 + i).standard_normal((n, d)).astype(np.float32)
        reqs.append(Request(coords=x, k=k, m=min(m, n), kernel_type=kt, lengthscale=ls,
                            predecessor=True, seed=i, tag=i))
    return reqs


def run_poisson(server: Server, reqs, rate_per_s: float, *, seed: int = 0):
    """Submit ``reqs`` at Poisson arrival times; return results in request order."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=len(reqs))
    futs = []
    t_start = time.perf_counter()
    t_next = 0.0
    for req, gap in zip(reqs, gaps):
        t_next += gap
        delay = t_start + t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append(server.submit(req))
    return [f.result() for f in futs]


def latency_summary(results) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    span = max(r.done for r in results) - min(r.arrival for r in results)
    q = sum(r.indices.shape[0] for r in results)
    return {"requests": len(results), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "requests_per_s": len(results) / span, "queries_per_s": q / span}


__all__ = ["Request", "Result", "Server", "make_stream", "run_poisson", "run_request",
           "latency_summary"]
