"""Benchmark-batch JSONL records (covertreex/telemetry/logs.py:126-311,
schemas.py:6-40) written from device batch telemetry; no GPU needed."""

import json

from covertreex_amd.telemetry import (REQUIRED_KEYS, SCHEMA_ID, BenchmarkLogWriter,
                                      batch_record_fields)

BATCH = {"start": 512, "size": 512, "leaders": 7, "survivors": 3, "orphans": 480,
         "far_levels": 12, "scan_evals": 9000, "mis_evals": 800, "nonleader_evals": 3000,
         "grid_edges": 250, "distance_edges": 40, "scan_ns": 2_000_000, "mis_ns": 500_000,
         "nonleader_ns": 250_000, "total_ns": 3_000_000, "conflict_ns": 400_000}


def test_records_carry_schema_identity(tmp_path):
    path = tmp_path / "log.jsonl"
    runtime = {"engine": "hip", "seeds": {"mis": 3}}
    with BenchmarkLogWriter(str(path), run_id="r1", runtime=runtime, metadata={"tag": "x"}) as w:
        w.record_batch(batch_index=1, batch_size=512, fields=batch_record_fields(BATCH))
        w.record_event(event="done", extra={"queries": 5})
    a, b = [json.loads(s) for s in path.read_text().splitlines()]
    assert all(k in a for k in REQUIRED_KEYS)
    assert a["schema_id"] == SCHEMA_ID == "covertreex.benchmark_batch.v2"
    assert a["schema_version"] == 2 and a["run_id"] == "r1"
    assert a["batch_event_index"] == 0 and b["batch_event_index"] == 1
    assert a["seed_pack"] == {"mis": 3} and a["runtime_engine"] == "hip" and a["tag"] == "x"
    assert a["selected"] == 7 and a["dominated"] == 505 and a["candidates"] == 512
    assert abs(a["traversal_ms"] - 2.0) < 1e-12 and abs(a["conflict_graph_ms"] - 0.4) < 1e-12
    assert abs(a["mis_ms"] - 0.5) < 1e-12
    assert a["conflict_adj_pairs"] == 290 and a["conflict_grid_local_edges"] == 250
    assert b["event"] == "done" and b["queries"] == 5 and b["conflict_pairwise_reused"] == 1


def test_run_hash_depends_on_runtime_and_seeds(tmp_path):
    def h(runtime):
        w = BenchmarkLogWriter(str(tmp_path / "h.jsonl"), runtime=runtime)
        w.close()
        return w._run_hash
    assert h({"a": 1, "seeds": {"s": 1}}) == h({"seeds": {"s": 1}, "a": 1})
    assert h({"a": 1, "seeds": {"s": 1}}) != h({"a": 1, "seeds": {"s": 2}})
    assert h({"a": 1}) != h({"a": 2})
