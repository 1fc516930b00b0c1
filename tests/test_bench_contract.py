"""bench.py's roofline inputs: the committed rocprofv3 traffic summary holds
both traversal kernels, and the per-bounce aggregation reads it."""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_traversal_traffic_from_committed_profile():
    d = json.loads(bench.TRAFFIC_FILE.read_text())
    t = bench.traversal_traffic(d)
    assert t is not None, "profile lacks a traversal kernel"
    ks = [d["kernels"][k] for k in bench.TRACE_KERNELS]
    assert t["hbm_bytes_per_dispatch"] == sum(k["hbm_bytes_per_dispatch"] for k in ks)
    assert min(k["l2_hit_rate"] for k in ks) <= t["l2_hit_rate"] <= max(k["l2_hit_rate"] for k in ks)
    assert set(t["valu_busy"]) == set(bench.TRACE_KERNELS)
    cal = d["copy_calibration"]
    assert cal["fetch_scale"] > 0 and cal["write_scale"] > 0


def test_traversal_traffic_absent():
    assert bench.traversal_traffic(None) is None
    assert bench.traversal_traffic({"kernels": {bench.TRACE_KERNELS[0]: {}}}) is None
