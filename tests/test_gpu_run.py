"""Device-resident orchestration (lpbf::DeviceRun through include/lpbf_run.h,
and the lpbf_gpu driver): per-step records, running maxima, snapshots and
output formats against the CPU oracle and the reference case0 files.

Tolerances: FP64 fields and peak temperatures 1e-9 K; melt widths from
molten rows exact; sub-grid widths (linear interpolation of the Tm crossing,
which divides by a temperature difference) 1e-6 um.  FP32: peak temperature
within 0.5 K of the FP64 oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import hedgehog_lpbf_amd as h

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "hedgehog_lpbf_amd", "lib", "lpbf_gpu")


def _case0():
    mat, pr, dom = h.Material(), h.apply_case("0", h.Process()), h.Domain()
    dt = h.dt_cfl(mat, dom)
    n = int(1.5e-3 / dt)
    k = h.make_const(mat, pr, dom, dt)
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    return mat, pr, dom, dt, n, k, plan, x_csv, wrapped


def _run(k, mat, pr, dom, plan, x_csv, wrapped, **kw):
    return h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan, x_csv=x_csv,
                       wrapped=wrapped, **kw)


def _same_records(ra, rb):
    """Equal step records; a batch replayed as a graph uses shorter bands (more
    workgroups), so only the FP64 energy sums may differ, in their last bits."""
    assert len(ra) == len(rb)
    for x, y in zip(ra, rb):
        assert (x.step, x.has_col, x.W_um, x.W_sub_um) == (y.step, y.has_col, y.W_um, y.W_sub_um)
        assert (x.stats.tmax, x.stats.arg, x.stats.ymin, x.stats.ymax, x.stats.molten) == \
            (y.stats.tmax, y.stats.arg, y.stats.ymin, y.stats.ymax, y.stats.molten)
        for f in ("sum_dT", "sum_src", "sum_loss"):
            assert abs(getattr(x.stats, f) - getattr(y.stats, f)) <= \
                1e-12 * (abs(getattr(x.stats, f)) + 1.0) * 240 * 120


def test_case0_device_run_fp64_matches_oracle(ora):
    print("\n[case] DeviceRun case0 FP64 batch 64", flush=True)
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    with _run(k, mat, pr, dom, plan, x_csv, wrapped, batch=64) as run:
        assert run.advance(n) == n
        recs = run.records()
        peak_T, peak_W, steady, done = run.summary()
        T = run.download()
    s = ora.defaults("0")
    T_o, pk, W, Ws, summ = ora.run(s, ora.dt_cfl(s), n)
    assert done == n and len(recs) == n
    assert np.abs(T - T_o).max() <= 1e-9
    tmax = np.array([r.stats.tmax for r in recs])
    assert np.abs(tmax - pk).max() <= 1e-9
    assert np.array_equal(np.array([r.W_um for r in recs]), W)
    cruise = (~wrapped) & (x_csv >= 0.45 * dom.Lx) & (x_csv <= 0.85 * dom.Lx)
    has = np.array([bool(r.has_col) for r in recs])
    assert np.array_equal(has, cruise)
    Wsub = np.array([r.W_sub_um for r in recs])
    assert cruise.sum() > 100
    assert np.abs(Wsub[cruise] - Ws[cruise]).max() <= 1e-6
    assert abs(peak_T - summ[0]) <= 1e-9 and peak_W == summ[1]
    assert abs(steady - summ[2]) <= 1e-6
    gold = json.load(open(os.path.join(GOLD, "case0_meta.json")))
    assert "%g" % peak_T == "%g" % gold["peak_T_K"]
    assert "%g" % peak_W == "%g" % gold["peak_W_um"]


@pytest.mark.parametrize("batch", [1, 7, 300])
def test_batch_size_does_not_change_results(batch):
    print(f"\n[case] DeviceRun case0 FP64 200 steps batch {batch}", flush=True)
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    n = 200
    plan, x_csv, wrapped = plan[:n], x_csv[:n], wrapped[:n]
    # shift the cruise window so it starts inside this short run
    x_csv = x_csv + 0.3 * dom.Lx
    with _run(k, mat, pr, dom, plan, x_csv, wrapped, batch=64) as a:
        a.advance(n)
        ra, Ta = a.records(), a.download()
    with _run(k, mat, pr, dom, plan, x_csv, wrapped, batch=batch) as b:
        got = 0
        while got < n:            # uneven advance calls across batch boundaries
            got += b.advance(min(37, n - got))
        rb, Tb = b.records(), b.download()
    assert np.array_equal(Ta, Tb)
    _same_records(ra, rb)


@pytest.mark.parametrize("phys", [0, 1])
def test_graph_batches_equal_eager_batches(phys):
    print(f"\n[case] DeviceRun case0 FP64 phys={phys}: hipGraph batches vs eager", flush=True)
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    if phys:
        dt = h.dt_cfl_ext(mat, dom, 1.0)
        k = h.make_const(mat, pr, dom, dt)
        plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    out = []
    for graph in (False, True):
        with _run(k, mat, pr, dom, plan, x_csv, wrapped, batch=64, phys=phys,
                  graph=graph) as r:
            r.advance(100)           # 64 aligned + 36: one graph batch, one eager
            r.advance(n - 100)       # unaligned start: eager until the next multiple of 64
            out.append((r.records(), r.download(), r.summary(), r.graph_launches()))
    (ra, Ta, sa, ga), (rb, Tb, sb, gb) = out
    assert ga == 0 and gb >= 15
    assert np.array_equal(Ta, Tb)
    _same_records(ra, rb)
    assert sa == sb


def test_dirichlet_noise_run_matches_oracle(ora):
    print("\n[case] DeviceRun 160x96 Dirichlet + noise FP64 300 steps", flush=True)
    mat, pr, dom = h.Material(k=7.0, rho=4430.0, cp=526.0), h.Process(P=200.0, V=1.0, r0=50e-6,
                                                                      h=50e-6), h.Domain()
    pr.Tamb = pr.Tsub = 300.0
    dom.nx, dom.ny, dom.Lx, dom.Ly = 160, 96, 159e-5, 95e-5
    dt = 5e-6
    k = h.make_const(mat, pr, dom, dt)
    n = 300
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    with _run(k, mat, pr, dom, plan, x_csv, wrapped, bc=h.LPBF_BC_DIRICHLET, T_bc=300.0,
              batch=50) as run:
        run.noise(0, 300.0, 0.5)
        T0 = run.download()
        run.advance(n)
        T = run.download()
        recs = run.records()
    assert np.array_equal(T0, ora.noise_field(0, 300.0, 0.5, (dom.ny, dom.nx)))
    T_o, st = ora.run_plan(k, plan, T0, 0, h.LPBF_BC_DIRICHLET, 300.0, mat.Tm)
    assert np.abs(T - T_o).max() <= 1e-9
    for r, o in zip(recs, st):
        assert abs(r.stats.tmax - o.tmax) <= 1e-9 and r.stats.arg == o.arg
        assert (r.stats.ymin, r.stats.ymax, r.stats.molten) == (o.ymin, o.ymax, o.molten)
        assert abs(r.stats.sum_dT - o.sum_dT) <= 1e-9 * dom.nx * dom.ny


def test_fp32_device_run_within_half_kelvin(ora):
    print("\n[case] DeviceRun case0 FP32 depth 4 batch 64", flush=True)
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    n = 600
    plan, x_csv, wrapped = plan[:n], x_csv[:n], wrapped[:n]
    dump = np.zeros(n, dtype=bool)
    dump[[0, 101, 333, n - 1]] = True
    with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan, x_csv=x_csv,
                     wrapped=wrapped, dump=dump, precision=32, depth=4, batch=64) as run:
        run.advance(n)
        recs = run.records()
        T = run.download()
        peak_T = run.summary()[0]
    T_o, st = ora.run_plan(k, plan, np.full((dom.ny, dom.nx), pr.Tamb))
    assert np.abs(T - T_o).max() <= 0.5
    tmax = np.array([r.stats.tmax for r in recs])
    assert np.abs(tmax - np.array([o.tmax for o in st])).max() <= 0.5
    assert abs(peak_T - max(o.tmax for o in st)) <= 0.5
    # hot columns only at launch ends: launches of 4 steps that stop at dump
    # steps and batch ends
    cruise = (~wrapped) & (x_csv >= 0.45 * dom.Lx) & (x_csv <= 0.85 * dom.Lx)
    ends = set()
    s = 0
    while s < n:
        b_end = (s // 64 + 1) * 64
        L = 1
        while L < 4 and s + L < min(b_end, n) and not dump[s + L - 1]:
            L += 1
        ends.add(s + L - 1)
        s += L
    has = np.array([bool(r.has_col) for r in recs])
    want = np.array([(i in ends) and cruise[i] for i in range(n)])
    assert np.array_equal(has, want)


def test_lpbf_gpu_driver_writes_reference_formats(tmp_path, ora):
    print("\n[case] lpbf_gpu case0 snap-every 200", flush=True)
    out = tmp_path / "case0"
    r = subprocess.run([EXE, "--case=0", "--snap-every=200", "--threads=3", f"--out={out}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert (out / "metrics.csv").read_text() == open(
        os.path.join(GOLD, "case0_metrics.csv")).read()
    meta = json.load(open(out / "meta.json"))
    gold = json.load(open(os.path.join(GOLD, "case0_meta.json")))
    for key in ("dt_s", "n_steps", "peak_T_K", "peak_W_um", "nx", "ny", "P_W", "V_mps", "r0_m"):
        assert "%g" % meta[key] == "%g" % gold[key], key
    snaps = sorted(p.name for p in out.glob("T_step_*.bin"))
    assert snaps == ["T_step_%06d.bin" % s for s in (0, 200, 400, 600, 800, 1000, 1200, 1400,
                                                     1426)]
    s = ora.defaults("0")
    T_o = ora.run(s, ora.dt_cfl(s), 1427)[0]
    T = np.fromfile(out / "T_step_001426.bin", dtype=np.float64).reshape(120, 240)
    assert np.abs(T - T_o).max() <= 1e-9


def test_lpbf_gpu_fp32_metrics_within_half_kelvin(tmp_path):
    print("\n[case] lpbf_gpu case0 --precision=32", flush=True)
    out = tmp_path / "case0_f32"
    r = subprocess.run([EXE, "--case=0", "--snap-every=200", "--precision=32", f"--out={out}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = [ln.split(",") for ln in (out / "metrics.csv").read_text().splitlines()[1:]]
    want = [ln.split(",") for ln in open(os.path.join(GOLD, "case0_metrics.csv")).read()
            .splitlines()[1:]]
    assert [g[0] for g in got] == [w[0] for w in want]          # the same dump steps
    for g, w in zip(got, want):
        assert g[1:3] == w[1:3]                                 # time and laser position
        assert abs(float(g[3]) - float(w[3])) <= 0.5 + 0.01     # peak T (printed to 6 digits)


def test_lpbf_gpu_manifest_and_quiet(tmp_path, ora):
    print("\n[case] lpbf_gpu --manifest (2 rows) and --quiet", flush=True)
    man = tmp_path / "runs.csv"
    man.write_text("run_id,cond_id,draw_id,config,P_W,V_mps,r0_m,eta,dsub_m,kmult,physics\n"
                   "a,c0,0,single,285,0.96,3.35e-05,0.35,3e-05,1,base\n"
                   "b,c1,1,single,325,1.2,3.35e-05,0.4,3e-05,1,base\n")
    r = subprocess.run([EXE, "--manifest=%s" % man, "--tend=4e-4"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [ln.split(",") for ln in r.stdout.splitlines() if ln.startswith("RESULT,")]
    assert [x[1] for x in rows] == ["a", "b"]
    for row, (P, V, eta) in zip(rows, [(285.0, 0.96, 0.35), (325.0, 1.2, 0.4)]):
        s = ora.defaults("0")
        s.proc.P, s.proc.V, s.mat.eta = P, V, eta
        dt = ora.dt_cfl(s)
        n = int(4e-4 / dt)
        summ = ora.run(s, dt, n)[4]
        assert int(row[15]) == n
        assert row[13] == "%.6g" % summ[0]          # peak_T_K
        assert row[12] == "%.6g" % summ[2]          # steady_W_sub_um
    # the same rows concurrently (one stream each): identical RESULT values, same order
    rj = subprocess.run([EXE, "--manifest=%s" % man, "--tend=4e-4", "--jobs=2"],
                        capture_output=True, text=True, timeout=300)
    assert rj.returncode == 0, rj.stderr
    rows_j = [ln.split(",") for ln in rj.stdout.splitlines() if ln.startswith("RESULT,")]
    assert [x[:14] + x[15:] for x in rows_j] == [x[:14] + x[15:] for x in rows]
    q = subprocess.run([EXE, "--case=0", "--quiet", "--tend=4e-4"], capture_output=True,
                       text=True, timeout=300)
    assert q.returncode == 0, q.stderr
    qrow = [ln for ln in q.stdout.splitlines() if ln.startswith("RESULT,")][0].split(",")
    assert qrow[12:14] == rows[0][12:14]


@pytest.mark.parametrize("f64_depth", [2, 4])
def test_blocked_fp64_run_equals_single_steps(f64_depth):
    """f64_depth > 1 runs stretches without cruise or dump steps on the
    temporally blocked FP64 kernel: fields bitwise, per-step records equal up
    to the last bits of the energy sums (other block partition), summaries
    equal; both physics, a 64-step batch and an odd one."""
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    n = 500
    for phys, batch, bc in ((0, 64, 0), (1, 37, 0), (0, 64, 1)):
        print(f"\n[case] DeviceRun case0 FP64 f64_depth={f64_depth} phys={phys} batch={batch} "
              f"bc={bc}", flush=True)
        kk, pl, xc, wr = k, plan[:n], x_csv[:n], wrapped[:n]
        if phys:
            dtx = h.dt_cfl_ext(mat, dom, 1.0)
            kk = h.make_const(mat, pr, dom, dtx)
            pl, xc, wr = h.single_track_plan(pr.V, dtx, dom.Lx, dom.Ly, n)
        out = []
        for fd in (1, f64_depth):
            with _run(kk, mat, pr, dom, pl, xc, wr, batch=batch, phys=phys, f64_depth=fd,
                      bc=bc, T_bc=300.0) as r:
                assert r.advance(n) == n
                out.append((r.records(), r.download(), r.summary()))
        (ra, Ta, sa), (rb, Tb, sb) = out
        assert np.array_equal(Ta, Tb)
        _same_records(ra, rb)
        assert sa == sb
        assert sum(1 for x in rb if x.has_col) > 0


def test_lpbf_gpu_f64_depth_writes_the_same_outputs(tmp_path):
    """Snapshots and metrics.csv of a blocked FP64 run equal the single-step
    run's byte for byte (launches end at every dump and cruise step)."""
    print("\n[case] lpbf_gpu case0 --f64-depth=3 --snap-every=7 vs --f64-depth=1", flush=True)
    dirs = []
    for fd in (1, 3):
        out = tmp_path / f"d{fd}"
        r = subprocess.run([EXE, "--case=0", "--tend=3e-4", "--snap-every=7", f"--f64-depth={fd}",
                            f"--out={out}"], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        dirs.append(out)
    a, b = dirs
    assert (a / "metrics.csv").read_bytes() == (b / "metrics.csv").read_bytes()
    snaps = sorted(p.name for p in a.glob("T_step_*.bin"))
    assert len(snaps) > 10 and snaps == sorted(p.name for p in b.glob("T_step_*.bin"))
    for name in snaps:
        assert (a / name).read_bytes() == (b / name).read_bytes(), name


def test_fp32_cruise_cols_sample_every_cruise_step(ora):
    """cruise_cols ends FP32 launches at cruise steps too: every cruise step
    gets its hot column, as in the FP64 path, and the steady sub-grid width
    is within 0.5 um of the FP64 run's (the FP32 field is within 0.5 K; the
    Tm crossing moves by dT / |dT/dy|, far below a cell at the melt edge)."""
    print("\n[case] DeviceRun case0 FP32 depth 4 cruise_cols vs FP64", flush=True)
    mat, pr, dom, dt, n, k, plan, x_csv, wrapped = _case0()
    cruise = (~wrapped.astype(bool)) & (x_csv >= 0.45 * dom.Lx) & (x_csv <= 0.85 * dom.Lx)
    with _run(k, mat, pr, dom, plan, x_csv, wrapped, precision=32, depth=4,
              cruise_cols=True) as r32:
        r32.advance(n)
        rec32, s32 = r32.records(), r32.summary()
    with _run(k, mat, pr, dom, plan, x_csv, wrapped) as r64:
        r64.advance(n)
        rec64, s64 = r64.records(), r64.summary()
    has = np.array([bool(x.has_col) for x in rec32])
    assert cruise.sum() > 100 and np.array_equal(has, cruise)
    w32 = np.array([x.W_sub_um for x in rec32])[cruise]
    w64 = np.array([x.W_sub_um for x in rec64])[cruise]
    assert np.abs(w32 - w64).max() <= 0.5
    assert abs(s32[2] - s64[2]) <= 0.5 and s64[2] > 10.0
    assert abs(s32[0] - s64[0]) <= 0.5


def test_fp32_full_plate_50k_steps_within_half_kelvin():
    """The FP32 accuracy target at its full size: the 8192^2 plate (dx 5 um,
    dt 1.2 us, 20-track raster, P 250 W, V 1.2 m/s), 50,000 steps from the
    same noisy start (300 K +/- 0.5 K), the FP32 kernel (depth 4, 46-row
    bands) against the blocked FP64 kernel (depth 2), which equals the
    oracle bit for bit wherever the oracle can run
    (tests/test_gpu_f64_blocked.py).  Field, every step's peak temperature
    and the run's peak within 0.5 K; the peak melt width, counted in whole
    rows, within one row."""
    print("\n[case] DeviceRun 8192^2 raster 50k steps FP32 depth 4 vs FP64 depth 2",
          flush=True)
    n, dx, dt = 8192, 5e-6, 1.2e-6
    mat = h.Material(k=7.0, rho=4430.0, cp=526.0, Tm=1609.0, eps=0.40, eta=0.35)
    pr = h.Process(P=250.0, V=1.2, r0=50e-6, h=50e-6, Tamb=300.0, Tsub=300.0)
    dom = h.Domain(Lx=(n - 1) * dx, Ly=(n - 1) * dx, nx=n, ny=n)
    steps = 50000
    r = h.centred_raster(dom.Lx, dom.Ly, pr.V, dt, steps, 20, 80e-6)
    plan = h.raster_plan(r["x0"], r["y0"], r["length"], r["hatch"], r["n_tracks"], r["V"],
                         dt, steps)
    k = h.make_const(mat, pr, dom, dt)
    out = {}
    for name, kw in (("f64", dict(f64_depth=2)), ("f32", dict(precision=32, depth=4, rows=46))):
        with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan,
                         **kw) as run:
            run.noise(0, pr.Tamb, 0.5)
            assert run.advance(steps) == steps
            tmax = np.array([x.stats.tmax for x in run.records()])
            out[name] = (run.download(), tmax, run.summary())
        print(f"[case] {name} done: peak T {out[name][2][0]:.3f} K", flush=True)
    (T64, t64, s64), (T32, t32, s32) = out["f64"], out["f32"]
    err = np.abs(T32 - T64).max()
    print(f"[case] max |T32 - T64| = {err:.3e} K over {steps} steps", flush=True)
    assert len(t64) == len(t32) == steps
    assert err <= 0.5
    assert np.abs(t32 - t64).max() <= 0.5
    assert abs(s32[0] - s64[0]) <= 0.5 and s64[0] > mat.Tm
    assert abs(s32[1] - s64[1]) <= dx * 1e6 + 1e-9
    assert T64.max() > 1000.0          # the last track is still hot at the end
