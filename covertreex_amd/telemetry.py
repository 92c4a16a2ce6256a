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

Device-build specifics, recorded under ``hip_*`` keys: ``conflict_graph_ms``
is the conflict pass (every far pair's distance, grid cells and conflict
bits) and ``mis_ms`` the near/far split, greedy MIS and leader linking, each
from the device wall clock; ``conflict_adj_pairs`` counts the conflict edges
(grid and distance) and ``hip_mis_evals`` the far pairs scored,
n_far * (n_far - 1) / 2 per level as the reference counts.

:func:`batch_payloads` rebuilds, from the device's conflict records
(``set_conflict_log``), the per-batch payload the reference's Rust path sends
to ``record_rust_batch`` (``emit_rust_batch``, src/lib.rs:1274-1371, filled
by ``batch_insert_collect``, src/algo/batch/mod.rs:802-956): parents and
levels of the batch's rows, the selected and dominated rows, and the conflict
CSR in batch-local indices.
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
        "conflict_graph_ms": ms(b["conflict_ns"]),
        "mis_ms": ms(b["mis_ns"]),
        "traversal_pairwise_ms": ms(b["scan_ns"]),
        "traversal_mask_ms": 0.0,
        "traversal_semisort_ms": 0.0,
        "conflict_pairwise_ms": ms(b["conflict_ns"]),
        "conflict_adj_pairs": int(b["grid_edges"]) + int(b["distance_edges"]),
        "conflict_grid_local_edges": int(b["grid_edges"]),
        "batch_order_strategy": batch_order,
        "traversal_engine": "hip",
        "engine": "hip",
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


def parse_conflict_log(words):
    """Split a conflict record (int32 words, the layout of
    ``cx_tree_conflict_log``) into per-batch lists of level entries
    ``(level, far ids, leader mask, covered mask, distance bits, grid bits)``
    (bool arrays; the bit matrices nf x nf, upper triangle) and bottom
    survivors."""
    import numpy as np

    w = np.asarray(words, np.int64)
    out: Dict[int, Dict[str, list]] = {}
    pos = 0

    def unpack(row_words, nf):
        u = (row_words.astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)
        bits = (u[..., :, None] >> np.arange(32, dtype=np.uint32)) & 1
        return bits.reshape(*u.shape[:-1], -1)[..., :nf].astype(bool)

    while pos < w.size:
        kind, batch = int(w[pos]), int(w[pos + 1])
        rec = out.setdefault(batch, {"levels": [], "survivors": []})
        if kind == 1:
            level, nf = int(w[pos + 2]), int(w[pos + 3])
            W = (nf + 31) // 32
            p = pos + 4
            far = w[p:p + nf]
            p += nf
            sel = unpack(w[p:p + W], nf)
            cov = unpack(w[p + W:p + 2 * W], nf)
            p += 2 * W
            dist = unpack(w[p:p + nf * W].reshape(nf, W), nf)
            grid = unpack(w[p + nf * W:p + 2 * nf * W].reshape(nf, W), nf)
            rec["levels"].append((level, far, sel, cov, dist, grid))
            pos = p + 2 * nf * W
        elif kind == 2:
            cnt = int(w[pos + 2])
            rec["survivors"].extend(int(x) for x in w[pos + 3:pos + 3 + cnt])
            pos += 3 + cnt
        else:
            raise ValueError(f"conflict record: unknown entry kind {kind} at word {pos}")
    return out


def batch_payloads(tree, *, coords=None, grid_scale: float = 1.0, words=None):
    """The reference's per-batch payload (``emit_rust_batch``) for every batch
    with a conflict record: ``batch_index``, ``batch_size``, ``parents``,
    ``levels`` (the batch's rows, tree order), ``selected`` (leaders at any
    level and bottom survivors) and ``dominated`` (far non-leaders covered by a
    leader), batch-local and ascending, ``conflict_indptr`` /
    ``conflict_indices`` (row u lists, level by level, its distance
    conflicts in far-list order, then each grid neighbour twice -- the
    reference pushes both directions of every directed grid edge; the order
    among grid neighbours follows a hash set there and ascending far-list
    order here), ``timings`` (device wall clock) and the grid statistics.
    ``coords`` (dataset order, as given to the build) adds
    ``conflict_grid_cells``: the most occupied grid cells of one level's far
    set."""
    return payloads_from_record(
        tree.conflict_log() if words is None else words, tree.build_telemetry(),
        tree.get_parents(), tree.get_levels(),
        order=tree.get_order() if coords is not None else None, coords=coords,
        grid_scale=grid_scale)


def _scopes_from_rows(rows, nb):
    """build_scopes_from_edges (src/algo/batch/mod.rs:16-200) with the
    reference's defaults (scope cap = batch length, no budget schedule, no
    segment limit, masked append, pair merge): each row's scope is its
    distinct conflict neighbours in ascending order, and the chunk
    statistics follow."""
    indptr, indices = [0], []
    max_members = points = scans = dedupe = pairs_after = 0
    for r in rows:
        members = sorted(set(r))
        scans += len(r)
        dedupe += len(r) - len(members)
        max_members = max(max_members, len(members))
        points += len(members)
        pairs_after += min(len(members) * (len(members) - 1) // 2, nb * nb)
        indices.extend(members)
        indptr.append(len(indices))
    pairs_before = sum(len(r) for r in rows) // 2
    pair_cap = max_members * max_members
    pairs_after = min(pairs_after, pair_cap)
    return {"scope_indptr": indptr, "scope_indices": indices,
            "scope_chunk_segments": nb, "scope_chunk_emitted": nb,
            "scope_chunk_max_members": max_members, "scope_chunk_points": points,
            "scope_chunk_scans": scans, "scope_chunk_dedupe": dedupe,
            "scope_chunk_saturated": 0, "scope_budget_start": nb, "scope_budget_final": nb,
            "scope_budget_escalations": 0, "scope_budget_early_terminate": 0,
            "scope_chunk_pair_cap": pair_cap, "scope_chunk_pairs_before": pairs_before,
            "scope_chunk_pairs_after": pairs_after,
            "scope_chunk_pair_merges": max(pairs_before - pairs_after, 0)}


def payloads_from_record(words, batches, parents, levels, *, order=None, coords=None,
                         grid_scale: float = 1.0):
    """:func:`batch_payloads` from its parts: the conflict record, the
    per-batch telemetry rows (``start``, ``size`` and, when present, the
    ``*_ns`` timings), the tree's parents and levels (tree order) and, for
    the grid statistics, the node-to-dataset order and dataset coordinates."""
    import numpy as np

    recs = parse_conflict_log(words)
    parents = np.asarray(parents)
    levels = np.asarray(levels)
    out = []
    for b, rec in sorted(recs.items()):
        st = batches[b]
        s0, nb = int(st["start"]), int(st["size"])
        rows = [[] for _ in range(nb)]
        sel = np.zeros(nb, bool)
        dom = np.zeros(nb, bool)
        cells = leaders_raw = local_edges = 0
        for level, far, lsel, lcov, dist, grid in rec["levels"]:
            loc = far - s0
            sel[loc[lsel]] = True
            dom[loc[lcov]] = True
            nf = far.size
            leaders_raw = max(leaders_raw, nf)
            local_edges += int(grid.sum())
            if coords is not None and nf:
                keys = np.floor(np.asarray(coords, np.float32)[order[far]] * np.float32(grid_scale))
                cells = max(cells, len({tuple(k) for k in keys.tolist()}))
            dsym = dist | dist.T
            gsym = grid | grid.T
            for i in range(nf):
                r = rows[int(loc[i])]
                r.extend(int(x) for x in loc[np.nonzero(dsym[i])[0]])
                for j in np.nonzero(gsym[i])[0]:
                    r.extend((int(loc[j]), int(loc[j])))
        for node in rec["survivors"]:
            sel[node - s0] = True
        indptr = np.zeros(nb + 1, np.int64)
        indptr[1:] = np.cumsum([len(r) for r in rows])
        indices = np.array([x for r in rows for x in r], np.int64)
        scopes = _scopes_from_rows(rows, nb)
        out.append({
            "batch_index": b,
            "batch_size": nb,
            "parents": parents[s0:s0 + nb].tolist(),
            "levels": levels[s0:s0 + nb].astype(np.int64).tolist(),
            "selected": np.nonzero(sel)[0].tolist(),
            "dominated": np.nonzero(dom)[0].tolist(),
            "conflict_indptr": indptr.tolist(),
            "conflict_indices": indices.tolist(),
            "scope_indptr": scopes.pop("scope_indptr"),
            "scope_indices": scopes.pop("scope_indices"),
            "timings": dict({"traversal_seconds": st.get("scan_ns", 0) * 1e-9,
                             "conflict_graph_seconds": st.get("conflict_ns", 0) * 1e-9,
                             "mis_seconds": st.get("mis_ns", 0) * 1e-9,
                             "pairwise_seconds": st.get("conflict_ns", 0) * 1e-9,
                             "mask_seconds": 0.0, "semisort_seconds": 0.0,
                             "degree_cap": 0, "degree_pruned_pairs": 0,
                             "mis_iterations": 1}, **scopes),
            "batch_order_strategy": "hilbert-morton",
            "engine": "hip",
            "conflict_grid_cells": cells if coords is not None else None,
            "conflict_grid_leaders_raw": leaders_raw,
            "conflict_grid_leaders_after": cells if coords is not None else None,
            "conflict_grid_local_edges": local_edges,
        })
    return out


__all__ = ["SCHEMA_ID", "SCHEMA_VERSION", "REQUIRED_KEYS", "BenchmarkLogWriter",
           "batch_record_fields", "record_build_batches", "parse_conflict_log",
           "batch_payloads", "payloads_from_record"]
