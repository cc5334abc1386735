"""Ensemble of runs on one plate (lpbf::Ensemble through include/lpbf_run.h,
and lpbf_gpu --manifest --ensemble) against independent DeviceRuns of the
same inputs.

Every cell update and hot-column copy of an ensemble run is the one a
DeviceRun makes, so fields, per-step integer statistics, tmax, W, W_sub and
the run summaries must be identical (bitwise).  The FP64 energy sums merge
bands of another height (the ensemble picks its own), so against DeviceRun
they are compared to 1e-12 relative; between two ensembles with the same
band height they are bitwise too."""
import os
import subprocess

import numpy as np
import pytest

import hedgehog_lpbf_amd as h

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hedgehog_lpbf_amd", "lib", "lpbf_gpu")

# (P, V, r0, eta, kmult): rows of a small process sweep
ROWS = [(285.0, 0.96, 33.5e-6, 0.35, 1.0), (325.0, 1.20, 33.5e-6, 0.40, 1.0),
        (245.0, 0.80, 41.0e-6, 0.30, 1.5), (285.0, 0.96, 24.5e-6, 0.35, 2.0),
        (300.0, 1.10, 30.0e-6, 0.45, 1.2)]


def _member(row, phys, t_end):
    P, V, r0, eta, kmult = row
    mat, pr, dom = h.Material(), h.Process(), h.Domain()
    pr.P, pr.V, pr.r0, mat.eta = P, V, r0, eta
    dt = h.dt_cfl_ext(mat, dom, kmult) if phys else h.dt_cfl(mat, dom)
    n = int(t_end / dt)
    k = h.make_const(mat, pr, dom, dt)
    k.kmult = kmult
    plan, x_csv, wrapped = h.single_track_plan(V, dt, dom.Lx, dom.Ly, n)
    return dict(k=k, plan=plan, x_csv=x_csv, wrapped=wrapped, Tamb=pr.Tamb), mat, dom


def _device_run(m, mat, dom, phys, **kw):
    return h.DeviceRun(m["k"], Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=m["Tamb"], plan=m["plan"],
                       x_csv=m["x_csv"], wrapped=m["wrapped"], phys=phys, **kw)


def _ints(r):
    return (r.step, r.has_col, r.W_um, r.W_sub_um, r.stats.tmax, r.stats.arg, r.stats.ymin,
            r.stats.ymax, r.stats.molten)


def _sums(r):
    return np.array([r.stats.sum_dT, r.stats.sum_src, r.stats.sum_loss])


@pytest.mark.parametrize("phys,bc", [(0, 0), (1, 0), (0, 1)])
def test_ensemble_equals_device_runs(phys, bc):
    t_end = 6e-4
    print(f"\n[case] Ensemble of {len(ROWS)} runs phys={phys} bc={bc} t_end={t_end} vs "
          "DeviceRuns", flush=True)
    built = [_member(r, phys, t_end) for r in ROWS]
    runs = [b[0] for b in built]
    mat, dom = built[0][1], built[0][2]
    lens = [len(m["plan"]) for m in runs]
    if phys:
        assert len(set(lens)) > 1          # kmult changes dt: runs of different length
    with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, phys=phys, batch=64, bc=bc,
                    T_bc=300.0) as ens:
        assert ens.advance(10 ** 6) == max(lens)
        assert ens.graph_launches() >= max(lens) // 64
        got = [(ens.download(i), ens.records(i), ens.summary(i)) for i in range(len(runs))]
    for i, m in enumerate(runs):
        with _device_run(m, mat, dom, phys, batch=64, bc=bc, T_bc=300.0) as dr:
            assert dr.advance(lens[i]) == lens[i]
            T, recs, summ = dr.download(), dr.records(), dr.summary()
        Te, re, se = got[i]
        assert se[3] == lens[i] and len(re) == lens[i]
        assert np.array_equal(Te, T), f"run {i}: field differs"
        assert [_ints(x) for x in re] == [_ints(y) for y in recs], f"run {i}: records differ"
        for x, y in zip(re, recs):
            assert np.all(np.abs(_sums(x) - _sums(y)) <= 1e-12 * (np.abs(_sums(y)) + 1.0) * 28800)
        assert se == summ, f"run {i}: summary differs"
        cruise = sum(1 for r in re if r.has_col)
        assert cruise > 0


def test_ensemble_chunking_graph_and_upload():
    print("\n[case] Ensemble: eager vs graph batches, odd advance chunks, uploaded field",
          flush=True)
    built = [_member(r, 0, 3e-4) for r in ROWS[:3]]
    runs = [b[0] for b in built]
    mat, dom = built[0][1], built[0][2]
    rng = np.random.default_rng(7)
    T0 = 300.0 + 400.0 * rng.random((dom.ny, dom.nx))
    out = []
    for graph, chunks in ((False, [37, 64, 10 ** 6]), (True, [64, 5, 64, 10 ** 6])):
        with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, batch=64, graph=graph) as ens:
            ens.upload(1, T0)
            for c in chunks:
                ens.advance(c)
            out.append(([ens.download(i) for i in range(3)],
                        [ens.records(i) for i in range(3)],
                        [ens.summary(i) for i in range(3)], ens.graph_launches()))
    (Ta, ra, sa, ga), (Tb, rb, sb, gb) = out
    assert ga == 0 and gb >= 2
    for i in range(3):
        assert np.array_equal(Ta[i], Tb[i])
        assert [_ints(x) for x in ra[i]] == [_ints(y) for y in rb[i]]
        assert all(np.array_equal(_sums(x), _sums(y)) for x, y in zip(ra[i], rb[i]))
    assert sa == sb
    # the uploaded start field against a DeviceRun started from it
    with _device_run(runs[1], mat, dom, 0, batch=64) as dr:
        dr.upload(T0)
        dr.advance(len(runs[1]["plan"]))
        assert np.array_equal(dr.download(), Tb[1])
        assert dr.summary() == sb[1]


def test_lpbf_gpu_manifest_ensemble_equals_row_loop(tmp_path):
    print("\n[case] lpbf_gpu --manifest (4 rows, base+ext) with and without --ensemble",
          flush=True)
    man = tmp_path / "runs.csv"
    man.write_text("run_id,cond_id,draw_id,config,P_W,V_mps,r0_m,eta,dsub_m,kmult,physics\n"
                   "a,c0,0,single,285,0.96,3.35e-05,0.35,3e-05,1,base\n"
                   "b,c1,1,single,325,1.2,3.35e-05,0.4,3e-05,1.5,ext\n"
                   "c,c2,2,single,245,0.8,4.1e-05,0.3,3e-05,1,base\n"
                   "d,c3,3,single,300,1.1,3e-05,0.45,3e-05,2,ext\n")
    res = []
    for extra in ([], ["--ensemble"]):
        r = subprocess.run([EXE, "--manifest=%s" % man, "--tend=4e-4"] + extra,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        res.append([ln.split(",") for ln in r.stdout.splitlines() if ln.startswith("RESULT,")])
    rows, rows_e = res
    assert [x[1] for x in rows_e] == ["a", "b", "c", "d"]
    # every column but wall_ms (index 14) identical
    assert [x[:14] + x[15:] for x in rows_e] == [x[:14] + x[15:] for x in rows]
    # the FP32 path runs baseline physics only: the ext rows are refused
    bad = subprocess.run([EXE, "--manifest=%s" % man, "--ensemble", "--precision=32"],
                         capture_output=True, text=True, timeout=120)
    assert bad.returncode == 2 and "baseline physics" in bad.stderr


@pytest.mark.parametrize("rows", [1, 3, 120])
def test_ensemble_band_heights_give_the_same_runs(rows):
    print(f"\n[case] Ensemble of 3 runs, band height {rows} vs the automatic one", flush=True)
    built = [_member(r, 1, 2e-4) for r in ROWS[:3]]
    runs = [b[0] for b in built]
    mat, dom = built[0][1], built[0][2]
    out = []
    for rw in (0, rows):
        with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, phys=1, batch=16,
                        rows=rw) as ens:
            ens.advance(10 ** 6)
            out.append(([ens.download(i) for i in range(3)], [ens.records(i) for i in range(3)]))
    (Ta, ra), (Tb, rb) = out
    for i in range(3):
        assert np.array_equal(Ta[i], Tb[i])
        assert [_ints(x) for x in ra[i]] == [_ints(y) for y in rb[i]]
        for x, y in zip(ra[i], rb[i]):
            assert np.all(np.abs(_sums(x) - _sums(y)) <= 1e-12 * (np.abs(_sums(y)) + 1.0) * 28800)
