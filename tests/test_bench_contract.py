"""bench.py's output contract on the GPU: one JSON line with the metric, the whole-job value, the
roofline object (bound, achieved, peak, unit, frac = achieved / peak, traffic from the committed
counter record when the kernel sources match it) and the workload config, on a short run of the
headline measurement (--quick: the DeepWalk-alias walk alone, two timed launches)."""
import json
import os
import subprocess
import sys

import pytest

from test_bench_checks import B as BENCH

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_prints_one_json_line_with_roofline():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--quick", "--no-cpu-baseline", "--steps", "2",
                        "--warmup", "1"], capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, lines
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["unit"] == "steps/s" and d["dtype"] == "f64" and d["vs_baseline"] is None
    steps = d["details"]["steps_per_launch"]
    assert steps == 79 * d["config"]["walkers"]
    assert abs(d["value"] - steps / (d["ms_per_step"] * 1e-3)) < 1e-6 * d["value"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["achieved"] - steps * 132 / (d["ms_per_step"] * 1e-3) / 1e9) < 1e-6 * rf["achieved"]
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12 and 0.0 < rf["frac"] < 1.0
    # Traffic comes from the committed counter record when it was taken on these kernel sources (and
    # is then within a few percent of the algorithmic bytes: no wasted re-reads), else it is null.
    with open(BENCH.TRAFFIC_FILE) as f:
        matches = json.load(f).get("source_sha256") == BENCH.kernel_source_hash()
    if matches:
        assert rf["traffic"] is not None and 0.95 < rf["traffic"] / (steps * 132) < 1.05
    else:
        assert rf["traffic"] is None
    assert d["config"]["workload"].startswith("deepwalk-alias")
    print("bench contract: %.3e steps/s, frac %.3f" % (d["value"], rf["frac"]), flush=True)
