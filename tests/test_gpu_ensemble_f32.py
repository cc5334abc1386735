"""FP32 ensemble (lpbf::Ensemble with precision 32, through include/lpbf_run.h
and lpbf_gpu --manifest --ensemble --precision=32) against FP32 DeviceRuns of
the same inputs, and against the FP64 ensemble.

The FP32 field does not depend on the launch depth or the band height, so
every ensemble field must equal the DeviceRun field bit for bit, and so must
the integer per-step statistics, tmax and W.  W_sub is taken at launch ends;
where the launch boundaries agree (equal run lengths and batch, or
cruise_cols, which ends launches at every cruise step on both sides) it is
identical too, and so are the run summaries.  The FP32 energy sums add
per-lane partials over bands of another height and are compared to the
batched kernel tests' tolerance (1e-5 K per cell)."""
import os
import subprocess

import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from tests.test_gpu_ensemble import ROWS, _device_run, _member

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hedgehog_lpbf_amd", "lib", "lpbf_gpu")


def _ints(r):
    return (r.step, r.stats.tmax, r.stats.arg, r.stats.ymin, r.stats.ymax,
            r.stats.molten, r.W_um)


def _full(r):
    return _ints(r) + (r.has_col, r.W_sub_um)


def _sums_close(x, y, cells):
    for f in ("sum_dT", "sum_src", "sum_loss"):
        assert abs(getattr(x.stats, f) - getattr(y.stats, f)) <= 1e-5 * cells, (x.step, f)


def _truncate(m, n):
    m = dict(m)
    for key in ("plan", "x_csv", "wrapped"):
        m[key] = np.asarray(m[key])[:n]
    return m


@pytest.mark.parametrize("bc", [0, 1])
def test_f32_ensemble_equals_device_runs(bc):
    t_end = 6e-4
    print(f"\n[case] FP32 Ensemble of {len(ROWS)} runs bc={bc} t_end={t_end} vs FP32 DeviceRuns",
          flush=True)
    built = [_member(r, 0, t_end) for r in ROWS]
    runs = [b[0] for b in built]
    mat, dom = built[0][1], built[0][2]
    lens = [len(m["plan"]) for m in runs]
    assert len(set(lens)) == 1           # baseline physics: one dt, equal lengths
    with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, batch=64, bc=bc, T_bc=300.0,
                    precision=32, depth=4) as ens:
        assert ens.advance(10 ** 6) == lens[0]
        got = [(ens.download(i), ens.records(i), ens.summary(i)) for i in range(len(runs))]
    for i, m in enumerate(runs):
        with _device_run(m, mat, dom, 0, batch=64, bc=bc, T_bc=300.0, precision=32,
                         depth=4) as dr:
            assert dr.advance(lens[i]) == lens[i]
            T, recs, summ = dr.download(), dr.records(), dr.summary()
        Te, re, se = got[i]
        assert se[3] == lens[i] and len(re) == lens[i]
        assert np.array_equal(Te, T), f"run {i}: field differs"
        assert [_full(x) for x in re] == [_full(y) for y in recs], f"run {i}: records differ"
        for x, y in zip(re, recs):
            _sums_close(x, y, dom.nx * dom.ny)
        assert se == summ, f"run {i}: summary differs"
        assert sum(1 for r in re if r.has_col) > 0


def test_f32_ensemble_runs_of_different_length_and_upload():
    print("\n[case] FP32 Ensemble: runs of 5 lengths (batch rebuilt at each end), depth 3, "
          "cruise columns, uploaded field, odd advance chunks", flush=True)
    built = [_member(r, 0, 6e-4) for r in ROWS]
    mat, dom = built[0][1], built[0][2]
    n = len(built[0][0]["plan"])
    lens = [n, n - 37, n - 100, n - 63, n - 1]
    runs = [_truncate(b[0], m) for b, m in zip(built, lens)]
    rng = np.random.default_rng(11)
    T0 = 300.0 + 400.0 * rng.random((dom.ny, dom.nx))
    kw = dict(batch=50, precision=32, depth=3, cruise_cols=True)
    with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, **kw) as ens:
        ens.upload(1, T0)
        done = 0
        for c in (37, 64, 10 ** 6):
            done += ens.advance(c)
        assert done == n
        got = [(ens.download(i), ens.records(i), ens.summary(i)) for i in range(len(runs))]
    for i, m in enumerate(runs):
        with _device_run(m, mat, dom, 0, **kw) as dr:
            if i == 1:
                dr.upload(T0)
            assert dr.advance(lens[i]) == lens[i]
            T, recs, summ = dr.download(), dr.records(), dr.summary()
        Te, re, se = got[i]
        assert se[3] == lens[i] and len(re) == lens[i]
        assert np.array_equal(Te, T), f"run {i}: field differs"
        assert [_full(x) for x in re] == [_full(y) for y in recs], f"run {i}: records differ"
        for x, y in zip(re, recs):
            _sums_close(x, y, dom.nx * dom.ny)
        assert se == summ, f"run {i}: summary differs"


def test_f32_ensemble_close_to_fp64_ensemble():
    print("\n[case] FP32 Ensemble vs FP64 Ensemble, 3 runs, t_end 4e-4", flush=True)
    built = [_member(r, 0, 4e-4) for r in ROWS[:3]]
    runs = [b[0] for b in built]
    mat, dom = built[0][1], built[0][2]
    out = []
    for prec in (64, 32):
        with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, batch=64,
                        precision=prec) as ens:
            ens.advance(10 ** 6)
            out.append(([ens.download(i) for i in range(3)], [ens.summary(i) for i in range(3)]))
    (T64, s64), (T32, s32) = out
    for i in range(3):
        err = float(np.abs(T32[i] - T64[i]).max())
        print(f"  run {i}: max|dT| FP32 vs FP64 = {err:.3e} K, peak T {s32[i][0]:.2f} / "
              f"{s64[i][0]:.2f} K", flush=True)
        assert err <= 0.05
        assert abs(s32[i][0] - s64[i][0]) <= 0.05


def test_lpbf_gpu_manifest_f32_ensemble_equals_f32_row_loop(tmp_path):
    print("\n[case] lpbf_gpu --manifest --precision=32 (4 base rows) with and without "
          "--ensemble", flush=True)
    man = tmp_path / "runs.csv"
    man.write_text("run_id,cond_id,draw_id,config,P_W,V_mps,r0_m,eta,dsub_m,kmult,physics\n"
                   "a,c0,0,single,285,0.96,3.35e-05,0.35,3e-05,1,base\n"
                   "b,c1,1,single,325,1.2,3.35e-05,0.4,3e-05,1,base\n"
                   "c,c2,2,single,245,0.8,4.1e-05,0.3,3e-05,1,base\n"
                   "d,c3,3,single,300,1.1,3e-05,0.45,3e-05,1,base\n")
    res = []
    for extra in ([], ["--ensemble"]):
        r = subprocess.run([EXE, "--manifest=%s" % man, "--tend=4e-4", "--precision=32",
                            "--depth=4"] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        res.append([ln.split(",") for ln in r.stdout.splitlines() if ln.startswith("RESULT,")])
    rows, rows_e = res
    assert [x[1] for x in rows_e] == ["a", "b", "c", "d"]
    # every column but wall_ms (index 14) identical
    assert [x[:14] + x[15:] for x in rows_e] == [x[:14] + x[15:] for x in rows]
