"""Batch telemetry as JSON lines in the reference's benchmark-batch schema.

The reference writes one record per batch insert through
``BenchmarkLogWriter.record_batch`` (covertreex/telemetry/logs.py:126-311,
schema id and required keys in covertreex/telemetry/schemas.py:6-40); its
Rust path fills the record from ``BatchInsertTelemetry`` via
``record_rust_batch`` (src/lib.rs:1274-1371).  :class:`BenchmarkLogWriter`
here writes records with the same identity fields (``schema_id``,
``schema_version``, ``run_id``, ``run_hash``, ``runtime_digest``,
``seed_pack``, ``batch_event_index``, ``timestamp``, ``runtime``,
``metadata``) and the same per-batch keys where the device build has the
quantity; :func:`record_build_batches` emits one record per device batch from
``HipTree.build_telemetry()``.

Device-build specifics, recorded under ``hip_*`` keys: the conflict pass
(every far pair's distance and conflict bit) and the greedy MIS run back to
back on the device and are timed as one figure (``conflict_graph_ms``;
``mis_ms`` is 0 and ``hip_conflict_mis_fused`` is 1); ``conflict_adj_pairs``
counts the conflict edges (grid and distance) and ``hip_mis_evals`` the
far pairs scored, n_far * (n_far - 1) / 2 per level as the reference counts.
"""

from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path
from typing import Any, Dict, Mapping

SCHEMA_VERSION = 2
SCHEMA_ID = f"covertreex.benchmark_batch.v{SCHEMA_VERSION}"
REQUIRED_KEYS = ("schema_id", "schema_version", "run_id", "run_hash", "timestamp",
                 "batch_event_index", "batch_index", "batch_size", "runtime", "metadata")


def _read_rss_bytes() -> int | None:
    try:
        with open("/proc/self/statm", "r", encoding="utf-8") as fh:
            pages = int(fh.read().split()[1])
        return pages * os.sysconf("SC_PAGE_SIZE")
    except (OSError, ValueError, IndexError):
        return None


def _hash_payload(payload: Mapping[str, Any]) -> str:
    text = json.dumps(payload, sort_keys=True, separators=(",", ":")).encode("utf-8")
    return hashlib.sha256(text).hexdigest()


class BenchmarkLogWriter:
    """Append-only JSONL writer with the benchmark-batch record identity."""

    def __init__(self, path: str, *, run_id: str | None = None,
                 runtime: Mapping[str, Any] | None = None,
                 metadata: Mapping[str, Any] | None = None, run_hash: str | None = None):
        self._path = Path(path).expanduser()
        self._path.parent.mkdir(parents=True, exist_ok=True)
        self._handle = self._path.open("a", encoding="utf-8")
        self._previous_rss = _read_rss_bytes()
        self._run_id = run_id or "run-0"
        self._runtime = dict(runtime or {})
        self._seed_pack = dict(self._runtime.get("seeds") or {})
        self._runtime_fields = {f"runtime_{k}": v for k, v in self._runtime.items()}
        self._metadata = dict(metadata or {})
        self._runtime_digest = _hash_payload({"runtime": self._runtime})
        self._run_hash = run_hash or _hash_payload(
            {"runtime_digest": self._runtime_digest, "seeds": self._seed_pack})
        self._event_index = 0

    @property
    def path(self) -> Path:
        return self._path

    def _base(self) -> Dict[str, Any]:
        return {"schema_id": SCHEMA_ID, "schema_version": SCHEMA_VERSION,
                "run_id": self._run_id, "run_hash": self._run_hash,
                "runtime_digest": self._runtime_digest, "seed_pack": self._seed_pack,
                "batch_event_index": int(self._event_index), "timestamp": time.time()}

    def _write(self, record: Dict[str, Any]) -> None:
        if self._runtime_fields:
            record.update(self._runtime_fields)
        if self._metadata:
            record.update(self._metadata)
        self._handle.write(json.dumps(record, sort_keys=True))
        self._handle.write("\n")
        self._handle.flush()
        self._event_index += 1

    def record_batch(self, *, batch_index: int, batch_size: int,
                     fields: Mapping[str, Any] | None = None,
                     extra: Mapping[str, Any] | None = None) -> None:
        rss = _read_rss_bytes()
        record = self._base()
        record.update({"batch_index": int(batch_index), "batch_size": int(batch_size),
                       "runtime": self._runtime, "metadata": self._metadata})
        record.update(dict(fields or {}))
        for key, value in dict(extra or {}).items():
            if value is not None:
                record[key] = value
        if rss is not None:
            record["rss_bytes"] = int(rss)
            if self._previous_rss is not None:
                record["rss_delta_bytes"] = int(rss - self._previous_rss)
        self._previous_rss = rss
        self._write(record)

    def record_event(self, *, event: str, extra: Mapping[str, Any] | None = None) -> None:
        record = self._base()
        record["event"] = str(event)
        record["runtime"] = self._runtime
        record.setdefault("conflict_pairwise_reused", 1)
        for key, value in dict(extra or {}).items():
            if value is not None:
                record[key] = value
        self._write(record)

    def has_records(self) -> bool:
        return self._event_index > 0

    def close(self) -> None:
        if self._handle and not self._handle.closed:
            self._handle.close()

    def __enter__(self) -> "BenchmarkLogWriter":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def batch_record_fields(b: Mapping[str, int], *, batch_order: str = "hilbert-morton") -> Dict:
    """Benchmark-batch keys for one device batch (``HipTree.build_telemetry()``
    entry)."""
    ms = lambda ns: float(ns) / 1e6  # noqa: E731
    size = int(b["size"])
    return {
        "candidates": size,
        "selected": int(b["leaders"]),
        "dominated": size - int(b["leaders"]),
        "mis_iterations": 1,
        "traversal_ms": ms(b["scan_ns"]),
        "conflict_graph_ms": ms(b["mis_ns"]),
        "mis_ms": 0.0,
        "traversal_pairwise_ms": ms(b["scan_ns"]),
        "traversal_mask_ms": 0.0,
        "traversal_semisort_ms": 0.0,
        "conflict_pairwise_ms": ms(b["mis_ns"]),
        "conflict_adj_pairs": int(b["grid_edges"]) + int(b["distance_edges"]),
        "conflict_grid_local_edges": int(b["grid_edges"]),
        "batch_order_strategy": batch_order,
        "traversal_engine": "hip",
        "engine": "hip",
        "hip_conflict_mis_fused": 1,
        "hip_batch_start": int(b["start"]),
        "hip_leaders": int(b["leaders"]),
        "hip_survivors": int(b["survivors"]),
        "hip_orphans": int(b["orphans"]),
        "hip_far_levels": int(b["far_levels"]),
        "hip_scan_evals": int(b["scan_evals"]),
        "hip_mis_evals": int(b["mis_evals"]),
        "hip_nonleader_evals": int(b["nonleader_evals"]),
        "hip_nonleader_ms": ms(b["nonleader_ns"]),
        "hip_batch_ms": ms(b["total_ns"]),
    }


def record_build_batches(writer: BenchmarkLogWriter, tree, *, first_batch: int = 0,
                         batch_order: str = "hilbert-morton") -> int:
    """One record per batch of ``tree`` (build and inserts) from index
    ``first_batch``; returns the number written."""
    batches = tree.build_telemetry()
    for i, b in enumerate(batches[first_batch:], start=first_batch):
        writer.record_batch(batch_index=i, batch_size=int(b["size"]),
                            fields=batch_record_fields(b, batch_order=batch_order))
    return len(batches) - first_batch


__all__ = ["SCHEMA_ID", "SCHEMA_VERSION", "REQUIRED_KEYS", "BenchmarkLogWriter",
           "batch_record_fields", "record_build_batches"]
