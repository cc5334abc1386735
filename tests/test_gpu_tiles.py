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
    # the reference driver's meta.json keys (src/main.cpp:428-467), which add
    # physics, kmult and steady_W_sub_um to the published case0 meta.json
    assert set(meta) == set(gold) | {"physics", "kmult", "steady_W_sub_um", "eta", "dsub_m",
                                     "hconv"}
    assert set(meta["phases_ms"]) == set(gold["phases_ms"])
    assert set(meta["task_ms"]) == set(gold["task_ms"])
    assert meta["eta"] == 0.35 and meta["hconv"] >= 0 and meta["dsub_m"] > 0
    phases = meta["phases_ms"]
    assert all(v >= 0 for v in phases.values())
    assert sum(phases.values()) <= meta["wall_ms"] + 5
    # perstep.csv: one row per step, the phase columns and their total
    with open(out / "perstep.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["step", "build_us", "push_us", "wait_us", "stitch_us", "advance_us",
                       "metrics_us", "snapshot_us", "total_us"]
    assert len(rows) == 1 + 1427 and [int(r[0]) for r in rows[1:]] == list(range(1427))
    for r in rows[1:]:
        v = [float(x) for x in r[1:]]
        assert min(v) >= 0 and sum(v[:-1]) <= v[-1] * 1.001 + 1.0
    tot = [sum(float(r[1 + q]) for r in rows[1:]) / 1e3 for q in range(7)]
    for q, name in enumerate(("build", "push", "wait", "stitch", "advance", "metrics",
                              "snapshot")):
        assert abs(tot[q] - phases[name]) <= 1e-3 * max(1.0, phases[name]), name
