"""The reference's seven process cases (src/main.cpp:97-109) and its extended
physics on the device-resident FP64 path, against the CPU oracle: final field
within 1e-9 K, per-step peak temperature within 1e-9 K, molten-row widths
exact, and the run maxima of the reference drivers (peak T, peak W, the
cruise-gated steady sub-grid width; widths to 1e-6 um)."""
import numpy as np
import pytest

import hedgehog_lpbf_amd as h

pytestmark = pytest.mark.gpu

CASES = [("0", "base", 1.0), ("1.1", "base", 1.0), ("1.2", "base", 1.0), ("2.1", "base", 1.0),
         ("2.2", "base", 1.0), ("3.1", "base", 1.0), ("3.2", "base", 1.0),
         ("0", "ext", 1.0), ("3.1", "ext", 2.5)]


@pytest.mark.parametrize("case,physics,kmult", CASES)
def test_case_matches_oracle(ora, case, physics, kmult):
    print(f"\n[case] DeviceRun case {case} {physics} kmult={kmult}, 700 steps", flush=True)
    phys = 1 if physics == "ext" else 0
    mat, pr, dom = h.Material(), h.apply_case(case, h.Process()), h.Domain()
    dt = h.dt_cfl_ext(mat, dom, kmult) if phys else h.dt_cfl(mat, dom)
    n = 700
    k = h.make_const(mat, pr, dom, dt)
    k.kmult = kmult
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan, x_csv=x_csv,
                     wrapped=wrapped, phys=phys, batch=64) as r:
        r.advance(n)
        T = r.download()
        recs = r.records()
        peak_T, peak_W, steady, _ = r.summary()
    s = ora.defaults(case)
    dt_o = ora.dt_cfl_ext(s, kmult) if phys else ora.dt_cfl(s)
    assert dt_o == dt
    T_o, pk, W, Ws, summ = ora.run(s, dt_o, n, phys=phys, kmult=kmult)
    assert np.abs(T - T_o).max() <= 1e-9
    assert np.abs(np.array([x.stats.tmax for x in recs]) - pk).max() <= 1e-9
    assert np.array_equal(np.array([x.W_um for x in recs]), W)
    assert abs(peak_T - summ[0]) <= 1e-9 and peak_W == summ[1]
    assert abs(steady - summ[2]) <= 1e-6
