"""bench.py end to end on the small H2 workload (32^3 cells, one timed step):
the one JSON line on stdout carries every field the benchmark contract names,
with the roofline and CPU-baseline objects, and the headline follows the rule
(guided only when within the bound and faster than the plain step)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line_on_h2():
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "h2", "--steps", "1", "--warmup", "1",
           "--side-steps", "1", "--cpu-cells", "512", "--ref-sample", "16"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=280, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 1 and d["dtype"] == "f64"
    assert d["unit"] == "cell-steps/s" and d["higher_is_better"] is True and d["value"] > 0
    assert d["vs_baseline"] is None and "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert 0.0 < rf["frac"] < 1.0 and rf["achieved"] == pytest.approx(rf["frac"] * rf["peak"])
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] == "port"
    s = d["surrogate"]
    guided = d["config"]["path"] == "surrogate_guided"
    assert guided == (s["meets_bound"] and s["faster_than_plain"])
    if guided:
        assert d["value"] == pytest.approx(s["value"]) and d["plain_value"] > 0
