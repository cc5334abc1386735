"""The step-task drop-in end to end: the reference CPU solver's host loop
(tiles with mirrored halos, a 4-thread task graph, stitch, laser law,
measure, dump rule) driving StepTileTask on the GPU must write the
reference's case0 metrics.csv rows and meta.json peaks."""
import csv
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "hedgehog_lpbf_amd", "lib", "lpbf_tiles")


def test_case0_through_step_task_graph(tmp_path):
    print("\n[case] lpbf_tiles case0 threads=4 tiles=8", flush=True)
    out = tmp_path / "case0"
    r = subprocess.run([EXE, "--case=0", "--threads=4", "--tiles=8", "--snap-every=200",
                        f"--out={out}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with open(out / "metrics.csv") as f:
        got = f.read()
    with open(os.path.join(GOLD, "case0_metrics.csv")) as f:
        want = f.read()
    assert got == want
    meta = json.load(open(out / "meta.json"))
    gold = json.load(open(os.path.join(GOLD, "case0_meta.json")))
    for key in ("dt_s", "n_steps", "peak_T_K", "peak_W_um", "nx", "ny"):
        assert "%g" % meta[key] == "%g" % gold[key], key
    assert meta["task_ms"]["n_calls"] == 8 * 1427
    assert os.path.getsize(out / "T_step_001426.bin") == 240 * 120 * 8
