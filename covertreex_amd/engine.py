"""Engine plug-in boundary with the ``hip`` engine.

Mirrors the reference's ``TreeEngine`` protocol, ``CoverTree`` wrapper,
registry and ``build_tree`` / ``get_engine`` (covertreex/engine.py:36-130,
473-523, 1137-1147).  ``HipEngine`` takes the place of ``RustHilbertEngine``
(covertreex/engine.py:920-1135): the build and every query run on the GPU
through ``libcovertreex_hip.so``.
"""

from __future__ import annotations

import os
import time
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, Mapping, Protocol

import numpy as np

from .native import HipTree
from .residual import ResidualBackend, build_residual_backend

DEFAULT_ENGINE = "hip"
SUPPORTED_ENGINES = ("hip",)
# kNN limits of the device kernels (the reference has none)
MAX_K = 256
MAX_D_MIXED = 16


@dataclass(frozen=True)
class RuntimeConfig:
    """The runtime knobs the hip engine reads (subset of the reference's
    RuntimeConfig, covertreex/runtime/config.py:52-244)."""

    metric: str = "residual_correlation"
    engine: str = DEFAULT_ENGINE
    batch_order_strategy: str = "hilbert"
    residual_grid_whiten_scale: float = 1.0
    knn_precision: str = "mixed"  # "mixed" (exact, certified) or "reference-f32"
    residual_seed: int = 0


def _format_knn_output(indices: np.ndarray, distances: np.ndarray, *, return_distances: bool):
    """Return shapes of the legacy PCCT knn (covertreex/engine.py:36-56)."""
    if not return_distances:
        if indices.shape[0] == 1:
            squeezed = indices[0]
            return squeezed if squeezed.shape[0] > 1 else squeezed[0]
        return indices
    if indices.shape[0] == 1:
        si, sd = indices[0], distances[0]
        if si.shape[0] == 1:
            return si[0], sd[0]
        return si, sd
    return indices, distances


@dataclass(frozen=True)
class CoverTree:
    """Engine-level tree wrapper delegating queries to its engine
    (covertreex/engine.py:59-94)."""

    engine: "TreeEngine"
    handle: Any
    metric: str
    backend: Any = None
    meta: Mapping[str, Any] = field(default_factory=dict)

    def knn(self, query_points: Any, *, k: int, return_distances: bool = False,
            predecessor_mode: bool = False, context: Any = None, runtime: Any = None) -> Any:
        return self.engine.knn(self, query_points, k=k, return_distances=return_distances,
                               predecessor_mode=predecessor_mode, context=context,
                               runtime=runtime)

    @property
    def build_seconds(self) -> float | None:
        v = self.meta.get("build_seconds")
        return float(v) if v is not None else None


class TreeEngine(Protocol):
    """Interface for pluggable cover-tree engines (covertreex/engine.py:97-130)."""

    name: str

    def build(self, points: np.ndarray, *, runtime: Any, context: Any = None,
              batch_size: int = 512, seed: int = 0, build_mode: str = "batch",
              log_writer: Any | None = None, scope_cap_recorder: Any | None = None,
              plan_callback: Callable | None = None, residual_backend: Any | None = None,
              residual_params: Mapping[str, Any] | None = None) -> CoverTree: ...

    def knn(self, tree: CoverTree, query_points: Any, *, k: int, return_distances: bool,
            predecessor_mode: bool = False, context: Any = None, runtime: Any = None) -> Any: ...


@dataclass
class HipHandle:
    tree: HipTree
    node_to_dataset: np.ndarray
    backend: ResidualBackend
    subtree_min_bounds: np.ndarray | None


def _query_telemetry_enabled() -> bool:
    """COVERTREEX_HIP_QUERY_TELEMETRY (or the reference's
    COVERTREEX_RUST_QUERY_TELEMETRY, src/lib.rs:404-406) set to 1/true."""
    for key in ("COVERTREEX_HIP_QUERY_TELEMETRY", "COVERTREEX_RUST_QUERY_TELEMETRY"):
        value = os.environ.get(key, "")
        if value == "1" or value.lower() == "true":
            return True
    return False


class HipEngine:
    """Residual-correlation cover tree built and queried on an MI355X."""

    name = "hip"

    def build(self, points: np.ndarray, *, runtime: RuntimeConfig | None = None,
              context: Any = None, batch_size: int = 512, seed: int = 0,
              build_mode: str = "batch", log_writer: Any | None = None,
              scope_cap_recorder: Any | None = None, plan_callback: Callable | None = None,
              residual_backend: ResidualBackend | None = None,
              residual_params: Mapping[str, Any] | None = None,
              compute_predecessor_bounds: bool = True) -> CoverTree:
        runtime = runtime or RuntimeConfig()
        if runtime.metric != "residual_correlation":
            raise ValueError("hip engine only supports the residual_correlation metric.")
        params = dict(residual_params or {})
        if residual_backend is None:
            residual_backend = build_residual_backend(
                np.asarray(points, dtype=np.float64),
                seed=int(params.get("seed", runtime.residual_seed)),
                inducing_count=int(params.get("inducing_count", 512)),
                variance=float(params.get("variance", 1.0)),
                lengthscale=params.get("lengthscale", 1.0),
                chunk_size=int(params.get("chunk_size", 512)),
                kernel_type=int(params.get("kernel_type", 0)))
        b = residual_backend
        variance = float(params.get("variance", b.variance))
        lengthscale = params.get("lengthscale", b.lengthscale)
        kernel_type = int(params.get("kernel_type", b.kernel_type))
        chunk = int(params.get("chunk_size", b.chunk_size or batch_size))
        ls = np.atleast_1d(np.asarray(lengthscale, dtype=np.float32))
        start = time.perf_counter()
        tree = HipTree.build_residual(b.v_matrix, b.p_diag, b.kernel_points_f32, variance, ls,
                                      chunk_size=chunk or batch_size,
                                      batch_order=runtime.batch_order_strategy,
                                      grid_whiten_scale=runtime.residual_grid_whiten_scale,
                                      kernel_type=kernel_type)
        build_seconds = time.perf_counter() - start
        if log_writer is not None:
            # one benchmark-batch record per device batch (src/lib.rs:1274-1371)
            from .telemetry import record_build_batches

            record_build_batches(log_writer, tree)
        n2d = tree.get_order()
        smin = tree.get_subtree_bounds()[0] if compute_predecessor_bounds else None
        handle = HipHandle(tree=tree, node_to_dataset=n2d, backend=b, subtree_min_bounds=smin)
        meta = {"build_seconds": build_seconds, "dataset_size": b.num_points,
                "node_to_dataset": n2d, "backend_dtype": "float32", "chunk_size": chunk,
                "batch_order": "hilbert-morton", "kernel_type": kernel_type,
                "predecessor_bounds_computed": compute_predecessor_bounds,
                "precision": runtime.knn_precision}
        return CoverTree(engine=self, handle=handle, metric=runtime.metric, backend=b, meta=meta)

    def knn(self, tree: CoverTree, query_points: Any, *, k: int, return_distances: bool,
            predecessor_mode: bool = False, context: Any = None,
            runtime: RuntimeConfig | None = None) -> Any:
        handle: HipHandle = tree.handle
        queries = np.asarray(query_points)
        if queries.ndim == 1:
            queries = queries.reshape(-1, 1)
        if queries.dtype.kind not in {"i", "u"}:  # covertreex/engine.py:1087-1092
            raise ValueError(
                "hip engine expects integer query payloads representing dataset indices.")
        q = np.asarray(queries, dtype=np.int64).reshape(-1)
        precision = (runtime.knn_precision if runtime is not None
                     else tree.meta.get("precision", "mixed"))
        # the device kernels keep a query's top-k in one wavefront's registers
        # (k <= 256) and the screen's coordinates in fixed arrays (d <= 16)
        if not 1 <= int(k) <= MAX_K:
            raise ValueError(f"hip engine supports 1 <= k <= {MAX_K}; got k={k}")
        if precision == "mixed" and handle.tree.info()["d"] > MAX_D_MIXED:
            raise ValueError(f"hip engine's mixed precision supports d <= {MAX_D_MIXED}; "
                             f"use precision='reference-f32' for d={handle.tree.info()['d']}")
        telemetry = _query_telemetry_enabled()
        handle.tree.set_query_telemetry(telemetry)
        idx, dist = handle.tree.knn_residual(q, int(k), predecessor_mode=predecessor_mode,
                                             mode=precision)
        if telemetry:
            # attach to the active op log, as the reference does with its
            # query telemetry (covertreex/engine.py:1115-1131)
            telem = handle.tree.last_query_telemetry()
            op_log = getattr(context, "op_log", None)
            if telem is not None and op_log is not None and hasattr(op_log, "add_metadata"):
                op_log.add_metadata(hip_query_telemetry=telem)
            handle.tree.clear_last_query_telemetry()
        return _format_knn_output(idx, dist, return_distances=return_distances)


_HIP_ENGINE = HipEngine()
_ENGINE_REGISTRY: Dict[str, TreeEngine] = {_HIP_ENGINE.name: _HIP_ENGINE}


def register_engine(engine: TreeEngine) -> None:
    _ENGINE_REGISTRY[engine.name] = engine


def get_engine(name: str) -> TreeEngine:
    engine = _ENGINE_REGISTRY.get(name)
    if engine is None:
        raise ValueError(f"Unknown engine '{name}'. Supported engines: "
                         f"{', '.join(sorted(_ENGINE_REGISTRY))}")
    return engine


def build_tree(points: np.ndarray, *, runtime: RuntimeConfig | None = None,
               engine: str | None = None, context: Any = None, batch_size: int = 512,
               seed: int = 0, build_mode: str = "batch", log_writer: Any | None = None,
               scope_cap_recorder: Any | None = None, plan_callback: Callable | None = None,
               residual_backend: Any | None = None,
               residual_params: Mapping[str, Any] | None = None,
               compute_predecessor_bounds: bool = True) -> CoverTree:
    """covertreex/engine.py:483-523."""
    runtime = runtime or RuntimeConfig()
    impl = get_engine(engine or runtime.engine or DEFAULT_ENGINE)
    kwargs: Dict[str, Any] = dict(runtime=runtime, context=context, batch_size=batch_size,
                                  seed=seed, build_mode=build_mode, log_writer=log_writer,
                                  scope_cap_recorder=scope_cap_recorder,
                                  plan_callback=plan_callback,
                                  residual_backend=residual_backend,
                                  residual_params=residual_params)
    if isinstance(impl, HipEngine):
        kwargs["compute_predecessor_bounds"] = compute_predecessor_bounds
    return impl.build(points, **kwargs)


__all__ = ["CoverTree", "TreeEngine", "HipEngine", "RuntimeConfig", "DEFAULT_ENGINE",
           "SUPPORTED_ENGINES", "build_tree", "get_engine", "register_engine"]
