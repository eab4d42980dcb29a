"""bench.py's output line: one JSON object with the fields the measurement
contract names, at a small size (two steps of 2000 Cu -> Cu histories each,
a 20-history CPU baseline)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_line_has_the_contract_fields():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--histories", "2000", "--steps", "2",
                        "--warmup", "1", "--variant-histories", "0", "--autotune-sample", "0",
                        "--cpu-histories", "20"], capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["dtype"] == "f64"
    assert d["config"]["workload"] == "c2_cu_cu_500keV" and d["config"]["histories_per_step"] == 2000
    assert d["config"]["histories"] == 4000
    assert d["value"] > 0 and abs(d["ms_per_step"] * 2e-3 * d["value"] - d["config"]["collisions"]) \
        <= 1e-6 * d["config"]["collisions"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] == "valu" and rf["launches"] == 2
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12 and 0 < rf["frac"] < 1
    # achieved = algorithmic flop-eq per launch over the average kernel time
    assert abs(rf["achieved"] * 1e12 * rf["avg_kernel_ms"] * 1e-3 - rf["flop_eq_per_launch"]) \
        <= 1e-9 * rf["flop_eq_per_launch"]
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]
    assert cb["cores"] == min(16, cb["os_cpu_count"])
