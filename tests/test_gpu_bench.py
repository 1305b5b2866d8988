"""The benchmark script's contract on the device, on short runs: one JSON line
with the metric, value, roofline and configuration keys (the headline two-rotor
workload over its first sectors, and the batching workload on a few blocks)."""

import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT, capture_output=True, text=True,
                         timeout=240)
    assert res.returncode == 0, res.stderr[-3000:]
    lines = [l for l in res.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, res.stdout
    return json.loads(lines[0])


def test_headline_bench_line(gpu_device):
    d = _bench("--max-sectors", "4", "--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--profile-every", "1",
               "--batching-blocks", "6")
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d
    assert d["value"] > 0 and d["n_gpus"] == 1 and d["dtype"] == "f64"
    assert d["config"]["n_channel"] == 800 and d["config"]["n_energy"] == 64 and d["config"]["n_sector"] == 4
    r = d["roofline"]
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and r["peak"] == 78.6
    assert r["achieved"] > 0 and 0 < r["frac"] < 1
    # the rate is the timed launches' flop over the union of their intervals,
    # which is at most their summed time and at least the longest launch
    busy = r["busy_ms"] * 1e-3
    assert abs(r["achieved"] - r["flop_per_launch"] * r["timed_launches"] / busy / 1e12) < 1e-6 * r["achieved"]
    assert r["avg_launch_us"] * 1e-3 <= r["busy_ms"] <= r["summed_launch_ms"] * (1 + 1e-9)
    assert r["mean_concurrent_launches"] >= 1 - 1e-9
    assert 0 < r["hbm_frac"] < 1
    assert 0 < d["sweep"]["frac_of_fp64_peak"] < 1
    assert d["parity"] is None  # a reduced run names no fixture comparison
    assert d["max_unitarity_defect"] < 1e-8
    m = d["mixed"]
    assert m["value"] > 0 and m["refine_steps"] == 2 and m["ratio_to_fp64"] > 0
    assert m["max_abs_dS_vs_fp64_run"] <= 1e-6 and m["max_unitarity_defect"] < 1e-8
    bt = d["batching"]
    assert bt["value"] > 0 and bt["finite"] and bt["max_abs_diff_vs_per_block"] == 0.0 and bt["n_buckets"] >= 1


def test_batching_bench_line(gpu_device):
    d = _bench("--workload", "batching", "--blocks", "6", "--max-sectors", "2", "--steps", "1", "--warmup", "1")
    assert d["value"] > 0 and d["config"]["n_block"] == 6 and d["config"]["n_sector"] == 2
    assert d["max_rel_diff_bucketed_vs_per_block"] == 0.0 and d["max_abs_diff_two_lanes_vs_one"] == 0.0
    assert sum(b["n_blocks"] for b in d["buckets"]) == 6
