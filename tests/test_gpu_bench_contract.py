"""bench.py's output contract: one JSON line with the metric, the whole-job
value, the timing fields, the roofline of the dominant kernel and the
oracle error of the timed product (run on the GPU at a few products)."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_bench_json_line(gpu):
    print("case bench.py --steps 3 --warmup 1 --no-cpu --no-gmres", flush=True)
    out = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--steps", "3", "--warmup", "1",
                          "--no-cpu", "--no-gmres"], capture_output=True, text=True, timeout=240, cwd=REPO)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "oracle_error", "fp64"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1 and d["higher_is_better"] is True
    n = d["config"]["blobs"]
    assert n == 32768 and "workload" in d["config"]
    assert abs(d["value"] - n * n / (d["ms_per_step"] * 1e-3)) <= 1e-6 * d["value"]
    rf = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf, key
    assert rf["peak"] == 157.3 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    # the pair kernel runs inside the product
    assert 0.0 < rf["kernel_mean_ms"] <= d["ms_per_step"] * 1.15
    assert 0.0 < rf["kernel_median_ms"] <= d["ms_per_step"] * 1.15 and rf["kernel_launches"] == 3
    assert d["oracle_error"]["rel_l2"] <= 1e-6 and d["fp64"]["oracle_error"]["rel_l2"] <= 1e-12
