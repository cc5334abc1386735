"""The benchmark's one JSON line, as the default run prints it (fewer steps):
the metric and unit, the whole-job value, the roofline of the dominant
kernel (HBM-bound, achieved / peak = frac, traffic from the counter
summary when the tiling matches) and the CPU baseline beside it."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def test_default_bench_line_keeps_the_contract():
    print("\n[case] bench.py --steps 80 --warmup 8 (default workload)", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "80",
                        "--warmup", "8"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["metric"] == d["unit"] == "cell-updates/s"
    assert d["n_gpus"] == 1 and d["steps"] == 80 and d["warmup"] == 8
    assert d["value"] > 1e11 and d["dtype"] == "f32" and "workload" in d["config"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-9
    assert rf["traffic"] is None or 0 < rf["traffic"] < rf["alg_bytes_per_launch"]
    cb = d["cpu_baseline"]
    assert cb["kind"] in ("port", "reference") and cb["cores"] >= 1 and cb["value"] > 0
    assert cb["unit"] == "cell-updates/s" and cb["sample"]
