"""The benchmark's one JSON line, as the default run prints it (fewer steps):
the metric and unit, the whole-job value from repeated windows of exactly
--steps steps, the roofline of the dominant kernel (the bound from its
counter summary, achieved = HBM bytes per launch / launch time, frac =
achieved / peak <= 1, the algorithmic-byte fraction under its own name),
the CPU baseline beside it, and the FP64 parity path on the same plate in
its own object with the same fields."""
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
    assert d["config"]["timed_windows"] >= 1
    for line, dtype in ((d, "f32"), (d["fp64"], "f64")):
        assert line["dtype"] == dtype and line["value"] > 1e11
        rf = line["roofline"]
        assert rf["bound"] in ("hbm", "valu") and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
        assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-9
        assert 0 < rf["frac"] <= 1.0, rf
        assert rf["traffic"] is None or 0 < rf["traffic"] < rf["alg_bytes_per_launch"]
        assert rf["alg_frac"] > 0 and rf["kernel_ms_per_launch"] > 0
        if rf["valu_frac"] is not None:
            assert 0 < rf["valu_frac"] <= 1.0
        cb = line["cpu_baseline"]
        assert cb["kind"] in ("port", "reference") and cb["cores"] >= 1 and cb["value"] > 0
        assert cb["unit"] == "cell-updates/s" and "step-task" in cb["sample"]
