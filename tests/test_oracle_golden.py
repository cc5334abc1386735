"""Pin the CPU oracle on the reference's own golden run (reference/case0).

The fixture files are the reference's data files copied verbatim
(tests/golden/case0_metrics.csv, case0_meta.json).  Values there are printed
with C++ ostream's default 6 significant digits, which Python's '%g'
reproduces, so each row is compared as text.
"""
import csv
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _golden_rows():
    with open(os.path.join(GOLD, "case0_metrics.csv")) as f:
        return list(csv.DictReader(f))


def _golden_meta():
    with open(os.path.join(GOLD, "case0_meta.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def case0(ora):
    s = ora.defaults("0")
    dt = ora.dt_cfl(s)
    n = int(1.5e-3 / dt)  # main.cpp:236 with the default --tend
    T, pk, W, Ws, summ = ora.run(s, dt, n)
    plan = ora.build_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n, 200)
    return dict(s=s, dt=dt, n=n, T=T, pk=pk, W=W, Ws=Ws, summ=summ, plan=plan)


def test_dt_and_step_count(case0):
    meta = _golden_meta()
    assert "%g" % case0["dt"] == "%g" % meta["dt_s"]
    assert case0["n"] == meta["n_steps"] == 1427


def test_metrics_rows_match_golden(case0):
    rows = _golden_rows()
    assert len(rows) == 9
    plan = case0["plan"]
    dumps = [s for s in range(case0["n"]) if plan[s].dump]
    assert dumps == [int(r["step"]) for r in rows]
    for r in rows:
        s = int(r["step"])
        got = {
            "t_us": "%g" % (plan[s].t * 1e6),
            "x_laser_um": "%g" % (plan[s].x_laser_csv * 1e6),
            "peakT_K": "%g" % case0["pk"][s],
            "W_um": "%g" % case0["W"][s],
        }
        for key, val in got.items():
            assert val == r[key], (s, key, val, r[key])


def test_run_summary_matches_meta(case0):
    meta = _golden_meta()
    assert "%g" % case0["summ"][0] == "%g" % meta["peak_T_K"]
    assert "%g" % case0["summ"][1] == "%g" % meta["peak_W_um"]


def test_laser_wraps_between_1000_and_1200(case0):
    plan = case0["plan"]
    wrap = [s for s in range(case0["n"]) if plan[s].wrapped and not plan[s - 1].wrapped]
    assert len(wrap) == 1 and 1000 < wrap[0] < 1200
    # after the wrap the logged x restarts from 0.2*Lx
    s = wrap[0]
    assert plan[s].x_laser_csv == 0.2 * case0["s"].dom.Lx


def test_tiled_run_equals_full_grid_run(ora, case0):
    """The reference CPU solver steps 8 row-strip tiles (meta.json ntiles=8);
    the oracle's full-grid form must give the same bits (lpbf_cuda.cu:14-18)."""
    s, dt = case0["s"], case0["dt"]
    T8, pk8, W8, _, summ8 = ora.run(s, dt, 300, ntiles=8)
    T1, pk1, W1, _, summ1 = ora.run(s, dt, 300, ntiles=0)
    assert np.array_equal(T8, T1)
    assert np.array_equal(pk8, pk1) and np.array_equal(W8, W1)
