"""bench.py's command-line contract, checked without a GPU: one GPU only
(--gpus above 1 exits with a message and status 2), and the workloads the
driver may ask for are all known to the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "8", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_every_workload_is_listed():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("plate8k_f32", "plate8k_f64", "batch256", "case0", "ensemble", "ensemble_f32",
              "run8k_f64", "run8k_f32", "surrogate", "surrogate_kernels", "plate8k_f32_ext",
              "copy"):
        assert w in r.stdout, w
