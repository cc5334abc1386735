"""Throughput of the z-averaged 2D FD heat step on one MI355X.

Workload (default, `--workload plate8k_f32`): an 8192 x 8192 plate, dx = dy
= 5 um, dt = 1.2 us, k = 7 W/mK, rho = 4430, cp = 526, h = 50 um, T0 = 300 K
+ seeded uniform noise of +/-0.5 K, 20-track bidirectional raster with
80 um hatch at P = 250 W, V = 1.2 m/s (the 50,000-step scan plan; the timed
steps are its first K), mirror edges, FP32 field (theta = T - 300 K) with the
fused per-step statistics (peak T + index, molten rows and count, energy
sums) in FP64 across the grid.  One bench step = one time step of the whole
plate.  `--workload plate8k_f64` runs the same plate on the FP64 parity path
(temporally blocked, bitwise the single-step field; depth and band height
autotuned);
`batch256` steps 256 process-parameter instances of uneven size together
(one bench step = one time step of every instance); `case0` times the
reference's published case on the device-resident FP64 path; `ensemble`
steps a 64-row manifest sweep of that plate together (lpbf::Ensemble) and
times the row-by-row loop beside it; `run8k_f64` / `run8k_f32` run the
device-resident run (per-step metrics) on the 8192 x 8192 plate.

Prints one JSON line: value = cell-updates/s over the timed steps, with the
inputs already resident in HBM; `roofline` for the dominant kernel (step
kernel duration from HIP events on its stream); `cpu_baseline` = the CPU
restatement of the reference step in the GPU path's precision (FP32 for the
FP32 plate, the FP64 oracle otherwise; all host threads on row ranges)
timed on a bounded sample of the same plate.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def plate_setup(nx=8192, ny=8192):
    import hedgehog_lpbf_amd as h
    mat = h.Material(k=7.0, rho=4430.0, cp=526.0, Tm=1609.0, eps=0.40, eta=0.35)
    pr = h.Process(P=250.0, V=1.2, r0=50e-6, h=50e-6, Tamb=300.0, Tsub=300.0)
    dx = 5e-6
    dom = h.Domain(Lx=(nx - 1) * dx, Ly=(ny - 1) * dx, nx=nx, ny=ny)
    dt = 1.2e-6
    n_plan = 50000
    r = h.centred_raster(dom.Lx, dom.Ly, pr.V, dt, n_plan, 20, 80e-6)
    plan = h.raster_plan(r["x0"], r["y0"], r["length"], r["hatch"], r["n_tracks"], r["V"],
                         dt, n_plan)
    k = h.make_const(mat, pr, dom, dt)
    return mat, pr, dom, dt, k, plan


def cpu_threads():
    """Host threads for the CPU baseline: the process's CPU affinity, at most 16
    (the host share of one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n or 1))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(k, plan, nx, ny, fp32=False, target_s=12.0):
    """The reference CPU solver's step path on the same plate, in the
    precision of the GPU line it stands beside: per step the driver thread
    builds 2 x threads row-strip tiles with mirrored halos, a pool of
    `threads` workers steps them, and the driver stitches them back
    (src/main.cpp:43-44, 125-155, 316-348, restated in C by the oracle,
    oracle/lpbf_steptask.c).  A first single step sizes the sample to about
    `target_s` seconds; measure() is left out (the GPU line's statistics are
    fused, the reference's measure is a separate serial scan)."""
    import numpy as np
    from oracle import pyoracle as ora
    T = 300.0 + ora.noise_field(0, 0.0, 0.5, (ny, nx))
    T = np.ascontiguousarray(T, dtype=np.float32 if fp32 else np.float64)
    kk = ora.StepConst.from_buffer_copy(bytes(k))
    nt = cpu_threads()
    ntiles = 2 * nt
    first = ora.steptask_run(T, kk, plan, 1, ntiles, nt)[2]
    steps = int(max(1, min(2000, target_s / max(first, 1e-6))))
    ph = ora.steptask_run(T, kk, plan[1:], steps, ntiles, nt)
    el = ph[2]
    return {"value": nx * ny * steps / el, "unit": "cell-updates/s", "cores": nt,
            "kind": "port",
            "sample": f"oracle {'FP32' if fp32 else 'FP64'} step-task path (C, -O3): "
                      f"{ntiles} row-strip tiles built, stepped by {nt} worker threads and "
                      f"stitched per step, {nx}x{ny} plate, {steps} steps, {el:.1f} s "
                      f"(build {ph[0]:.1f} s, wait+stitch {ph[1]:.1f} s); host "
                      f"'{cpu_model()}', nproc {os.cpu_count()}, affinity "
                      f"{len(os.sched_getaffinity(0))}, threads capped at 16 (one GPU's "
                      f"host share)"}


def counters_for(workload, **launch):
    """The committed counter summary (profiles/counters_<workload>.json) when
    it was measured on the same launch shape (depth and band rows): HBM bytes
    per launch (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), VALU
    instructions per launch and the bound they show."""
    p = os.path.join(ROOT, "profiles", f"counters_{workload}.json")
    if os.path.exists(p):
        with open(p) as f:
            t = json.load(f)
        for sh in t.get("shapes", [t]):     # one launch shape, or a list of them
            if all(sh.get(k) == v for k, v in launch.items()):
                return dict(sh, summary=t.get("summary"))
    return None


def batch_setup(n_inst=256, n_steps=5000, seed=0):
    """256 instances, sides log-uniform in [256, 2048] per axis, P ~ U[100, 400] W,
    V ~ U[0.5, 2] m/s, r0 ~ U[35, 70] um, dx = 10 um, dt = 0.8 * dx^2 / (4 alpha),
    single track at mid-height with the reference wrap law, T0 = 300 K."""
    import numpy as np

    import hedgehog_lpbf_amd as h
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(256), np.log(2048), size=(n_inst, 2)))).astype(int)
    P = rng.uniform(100.0, 400.0, n_inst)
    V = rng.uniform(0.5, 2.0, n_inst)
    r0 = rng.uniform(35e-6, 70e-6, n_inst)
    mat = h.Material(k=7.0, rho=4430.0, cp=526.0, Tm=1609.0, eps=0.40, eta=0.35)
    dx = 1e-5
    dt = 0.8 * dx * dx / (4.0 * mat.alpha())
    specs = []
    for i in range(n_inst):
        nx, ny = int(sides[i, 0]), int(sides[i, 1])
        pr = h.Process(P=float(P[i]), V=float(V[i]), r0=float(r0[i]), h=50e-6, Tamb=300.0,
                       Tsub=300.0)
        dom = h.Domain(Lx=(nx - 1) * dx, Ly=(ny - 1) * dx, nx=nx, ny=ny)
        k = h.make_const(mat, pr, dom, dt)
        plan = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n_steps)[0]
        cells = nx * ny
        cls = 0 if cells < (1 << 18) else (1 if cells < (1 << 20) else 2)
        specs.append(h.InstanceSpec(k=k, plan=plan, T0=300.0, T_ref=300.0, T_bc=300.0,
                                    Tm=mat.Tm, size_class=cls))
    return specs, dt


def batch_bench(args):
    """All instances advance together: the autotuner picks depth and band height
    per size class, the warm-up runs eagerly, the timed steps are one hipGraph
    launch of the whole schedule."""
    import numpy as np
    import torch

    import hedgehog_lpbf_amd as h
    from oracle import pyoracle as ora
    steps = args.steps if args.steps is not None else 5000
    warmup = args.warmup if args.warmup is not None else 200
    specs, dt = batch_setup(n_steps=max(steps, warmup, 4))
    cells_total = sum(sp.k.nx * sp.k.ny for sp in specs)
    print(f"[bench] batch256: {len(specs)} instances, {cells_total} cells, dt {dt:.4g} s, "
          f"warmup={warmup} steps={steps}", file=sys.stderr, flush=True)
    b = h.F32Batch(specs, bc=0)
    s = torch.cuda.Stream()
    t0 = time.perf_counter()
    if args.batch_tiling:
        # fixed tiling per size class, "depth:rows" for classes 0, 1, 2
        for c, dr in enumerate(args.batch_tiling.split(",")):
            d, r = (int(v) for v in dr.split(":"))
            b.set_tiling(c, d, r)
        tiling = {c: b.tiling(c) for c in (0, 1, 2)}
        print(f"[bench] fixed tiling {tiling}", file=sys.stderr, flush=True)
    else:
        ms_tuned = b.autotune(4, s.cuda_stream)
        tiling = {c: b.tiling(c) for c in (0, 1, 2)}
        print(f"[bench] autotune {time.perf_counter() - t0:.1f} s: {tiling} "
              f"({ms_tuned:.3f} ms/step for the schedule)", file=sys.stderr, flush=True)
    with torch.cuda.stream(s):
        b.enqueue(0, warmup, s.cuda_stream)
        b.capture(0, steps, s.cuda_stream)
    torch.cuda.synchronize()
    h.check()
    n_launch, cells = b.launches(steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    with torch.cuda.stream(s):
        b.launch_graph(s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    h.check()
    total_ms = e0.elapsed_time(e1)
    value = cells * steps / (total_ms * 1e-3)
    peak = max(st.tmax for st in b.stats(0, 0, steps))
    print(f"[bench] graph of {n_launch} step launches: {total_ms:.1f} ms, instance 0 peak T "
          f"{peak:.1f} K", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        from concurrent.futures import ThreadPoolExecutor
        nt = cpu_threads()
        done = 0
        n_inst = 0
        with ThreadPoolExecutor(nt) as pool:
            t0 = time.perf_counter()
            for sp in specs:
                kk = ora.StepConst.from_buffer_copy(bytes(sp.k))
                T = np.full((sp.k.ny, sp.k.nx), 300.0, dtype=np.float32)
                out = np.empty_like(T)
                for q in range(40):
                    kk.x_l, kk.y_l, kk.q_scale = sp.plan[q]
                    ora.step_threads_f32(T, kk, threads=nt, out=out, pool=pool)
                    T, out = out, T
                done += 40 * sp.k.nx * sp.k.ny
                n_inst += 1
                if time.perf_counter() - t0 > 12.0:
                    break
            el = time.perf_counter() - t0
        cpu = {"value": done / el, "unit": "cell-updates/s", "cores": nt, "kind": "port",
               "sample": f"oracle FP32 step (C, -O3) on {nt} host threads, first {n_inst} "
                         f"instances x 40 steps, {el:.1f} s"}
    # roofline: the step kernels are issue-bound like the single plate's.
    # HBM bytes and VALU instructions per cell-update come from the committed
    # counter summary (every step launch of one graph replay, summed; they
    # vary by a few percent with the band heights the autotuner picks), else
    # the compulsory bytes: one read and one write of every instance per
    # launch.
    prof = counters_for("batch256")
    rate = cells * steps / (total_ms * 1e-3)            # cell-updates/s
    depth_of = {c: d for c, (d, r, m) in tiling.items()}
    compulsory = sum(8.0 * sp.k.nx * sp.k.ny / depth_of[sp.size_class] for sp in specs)
    if prof:
        bytes_step = prof["hbm_bytes_per_cell_update"] * cells
        basis = (f"counters (FETCH_SIZE + WRITE_SIZE, copy-calibrated) per cell-update, "
                 f"measured at tiling {prof['tiling']}")
        valu_frac = prof["valu_wave_insts_per_cell_update"] * rate / VALU_ISSUE_PER_S
    else:
        bytes_step, basis, valu_frac = compulsory, "compulsory bytes (8 B x cells / depth)", None
    achieved = bytes_step * steps / (total_ms * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": total_ms / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(specs)} instances, sides log-uniform 256-2048, P U[100,400] W, "
                        f"V U[0.5,2] m/s, r0 U[35,70] um, dx 10 um, dt 0.8x stability, "
                        f"mirror edges, FP32, per-step stats, one hipGraph launch",
            "cells_per_step": cells, "instances": len(specs), "graph_launches": n_launch,
            "tiling_per_class": {str(c): {"depth": d, "rows": r, "members": m}
                                 for c, (d, r, m) in tiling.items()},
        },
        "roofline": {
            "bound": prof.get("bound", "valu") if prof else "valu",
            "kernel": "k_step_f32_multi (whole graph incl. stats merges)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": bytes_step,
            "traffic_basis": basis, "traffic_unit": "bytes per step of all instances",
            "valu_frac": valu_frac,
            "alg_frac": 8.0 * rate / (HBM_PEAK_GBS * 1e9),
            "counter_summary": prof.get("summary") if prof else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    b.close()


def run8k_bench(args, precision):
    """The device-resident run (lpbf::DeviceRun, the reference GpuStepTask +
    MetricsTask pair) on the 8192^2 plate: per-step statistics rings,
    per-step metrics finalised on the host, one synchronisation per 64-step
    batch.  FP64: the steps on the temporally blocked FP64 kernel (--depth,
    default 2; 1 = single-step kernels); FP32: the FP32 kernel (--depth,
    default 4).  The raster plan has no cruise gate, so no hot columns are
    taken."""
    import torch

    import hedgehog_lpbf_amd as h
    mat, pr, dom, dt, k, plan = plate_setup(args.n, args.n)
    depth = args.depth if args.depth > 0 else (2 if precision == 64 else 4)
    rows = args.rows if args.rows > 0 or precision == 64 else 92
    warm = args.warmup if args.warmup is not None else 64
    steps = args.steps if args.steps is not None else 1024
    kw = dict(f64_depth=depth) if precision == 64 else dict(precision=32, depth=depth)
    with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb,
                     plan=plan[:warm + steps], batch=64, keep_records=False, rows=rows,
                     **kw) as r:
        r.noise(0, pr.Tamb, 0.5)
        r.advance(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.advance(steps)
        wall = time.perf_counter() - t0
        peak_T = r.summary()[0]
    value = dom.nx * dom.ny * steps / wall
    print(json.dumps({
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == 64 else "f32", "data": "synthetic",
        "config": {"workload": f"{dom.nx}x{dom.ny} plate, 20-track raster P 250 W V 1.2 m/s, "
                               f"FP{precision} device-resident run (per-step statistics and "
                               f"metrics, batch 64), "
                               + (f"blocked FP64 kernel depth {depth}" if precision == 64
                                  else f"FP32 kernel depth {depth}"),
                   "depth": depth, "band_rows": rows, "peak_T_K": peak_T},
        "roofline": None, "cpu_baseline": None}), flush=True)


def case0_bench(args):
    """The reference's own published run (case0, 240x120, 1427 steps, FP64) on the
    device-resident path: field on the GPU, one host sync per 64-step batch,
    per-step metrics as the reference drivers compute them."""
    import torch

    import hedgehog_lpbf_amd as h
    mat, pr, dom = h.Material(), h.apply_case("0", h.Process()), h.Domain()
    dt = h.dt_cfl(mat, dom)
    n = int(1.5e-3 / dt)
    k = h.make_const(mat, pr, dom, dt)
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    reps = args.steps if args.steps is not None else 20
    for _ in range(args.warmup if args.warmup is not None else 2):
        with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan,
                         x_csv=x_csv, wrapped=wrapped, batch=64) as r:
            r.advance(n)
    torch.cuda.synchronize()
    walls = []
    for _ in range(reps):
        with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan,
                         x_csv=x_csv, wrapped=wrapped, batch=64) as r:
            t0 = time.perf_counter()
            r.advance(n)
            walls.append(time.perf_counter() - t0)
            peak_T = r.summary()[0]
    wall = sorted(walls)[len(walls) // 2]
    value = dom.nx * dom.ny * n / wall
    ref_value = dom.nx * dom.ny * n / 0.198      # reference/case0/meta.json wall_ms 198
    print(json.dumps({
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": n, "warmup": 0, "ms_per_step": wall * 1e3 / n, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": value / ref_value, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "reference case0: 240x120, 1427 steps, FP64 parity kernels, "
                               "device-resident run, batch 64, per-step metrics",
                   "median_of": reps, "peak_T_K": peak_T,
                   "reference_cpu_wall_ms": 198.0},
        "roofline": None, "cpu_baseline": None}), flush=True)


def ensemble_bench(args, precision=64):
    """The reference's manifest sweep shape: `--runs` rows of the case0 plate
    (240x120, baseline physics, t_end 1.5 ms) with P, V, r0, eta drawn from the
    reference case table's ranges, stepped together by lpbf::Ensemble (FP64:
    one launch per step for every row; FP32: the batched kernel, up to 4 steps
    per launch for every row).  value = cell-updates of all rows / wall; the
    same rows one after another on DeviceRun of the same precision (the row
    loop) are timed beside it."""
    import numpy as np
    import torch

    import hedgehog_lpbf_amd as h
    rng = np.random.default_rng(3)
    runs, t_end = [], 1.5e-3
    for _ in range(args.runs):
        mat, pr, dom = h.Material(), h.Process(), h.Domain()
        pr.P, pr.V = rng.uniform(245.0, 325.0), rng.uniform(0.8, 1.2)
        pr.r0, mat.eta = rng.uniform(24.5e-6, 41e-6), rng.uniform(0.30, 0.45)
        dt = h.dt_cfl(mat, dom)
        n = int(t_end / dt)
        k = h.make_const(mat, pr, dom, dt)
        plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
        runs.append(dict(k=k, plan=plan, x_csv=x_csv, wrapped=wrapped, Tamb=pr.Tamb))
    cells = dom.nx * dom.ny * sum(len(r["plan"]) for r in runs)
    reps = args.steps if args.steps is not None else 10

    pk = dict(precision=precision, depth=4) if precision == 32 else {}

    def ens_once(**kw):
        with h.Ensemble(runs, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, batch=64,
                        keep_records=False, **(kw or pk)) as e:
            t0 = time.perf_counter()
            e.advance(10 ** 9)
            wall = time.perf_counter() - t0
            return wall, [e.summary(i)[0] for i in range(len(runs))]

    def loop_once():
        t0 = time.perf_counter()
        peaks = []
        for r in runs:
            with h.DeviceRun(r["k"], Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=r["Tamb"],
                             plan=r["plan"], x_csv=r["x_csv"], wrapped=r["wrapped"], batch=64,
                             keep_records=False, **pk) as d:
                d.advance(len(r["plan"]))
                peaks.append(d.summary()[0])
        return time.perf_counter() - t0, peaks

    for _ in range(args.warmup if args.warmup is not None else 2):
        ens_once()
    torch.cuda.synchronize()
    ens = [ens_once() for _ in range(reps)]
    wall = sorted(w for w, _ in ens)[len(ens) // 2]
    loop = [loop_once() for _ in range(max(1, reps // 5))]
    lwall = sorted(w for w, _ in loop)[len(loop) // 2]
    assert ens[0][1] == loop[0][1], "ensemble and row loop disagree on peak T"
    extra = {}
    if precision == 32:   # the FP32 rows against the FP64 parity ensemble
        p64 = ens_once(precision=64)[1]
        extra["max_peak_T_diff_vs_fp64_K"] = max(abs(a - b) for a, b in zip(ens[0][1], p64))
    n_steps = max(len(r["plan"]) for r in runs)
    print(json.dumps({
        "metric": "cell-updates/s", "value": cells / wall, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": n_steps, "warmup": 0, "ms_per_step": wall * 1e3 / n_steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == 64 else "f32", "data": "synthetic",
        "config": {"workload": f"manifest sweep: {len(runs)} rows of the case0 plate (240x120, "
                               "baseline physics, 1.5 ms, P U[245,325] W, V U[0.8,1.2] m/s, "
                               "r0 U[24.5,41] um, eta U[0.30,0.45]) stepped together by "
                               "lpbf::Ensemble, " + ("FP64 parity kernels, batch 64"
                                                     if precision == 64 else
                                                     "batched FP32 kernel, depth 4, batch 64"),
                   **extra,
                   "median_of": reps, "rows": len(runs),
                   "row_loop_cell_updates_per_s": cells / lwall,
                   "row_loop_wall_ms": lwall * 1e3, "ensemble_wall_ms": wall * 1e3,
                   "speedup_vs_row_loop": lwall / wall},
        "roofline": None, "cpu_baseline": None}), flush=True)


def surrogate_bench(args):
    """The learned melt-pool window in the two-level run (the surrogate
    workload): training data from FD runs (64 seeded 768^2 runs, pairs every
    16 steps), training with the library's BF16-MFMA forward/backward kernels
    and Adam (batch 256, lr 1e-3 decaying to 5e-5), the held-out error, then a
    4096^2 plate (10-track raster, P 300 W, V 1 m/s, 40,000 steps) run three
    ways and each compared with the FP64 path: the plain FP32 run of the
    whole plate, the two-level run with the FD window, and the two-level run
    with the network (recheck every 8 calls, fallback above 2 K).  value =
    the learned run's simulated seconds per wall second."""
    import numpy as np
    import torch

    from hedgehog_lpbf_amd import surrogate as sg
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    steps = args.steps if args.steps is not None else 40000
    if steps % 16:
        raise SystemExit("--steps must be a multiple of 16 for the surrogate workload")
    train_info = None
    if args.load_params:
        model = sg.load_params(args.load_params)
        log(f"[bench] surrogate parameters from {args.load_params}")
    else:
        spec = sg.DataSpec(n_runs=64, nx=768, ny=768, steps=args.data_steps, seed=0)
        xs, Ds, tg = sg.generate(spec, log=log)
        del Ds
        N = xs.shape[0]
        rng = np.random.default_rng(0)
        perm = rng.permutation(N)
        n_test = max(N // 50, 512)
        test_idx, train_idx = perm[:n_test], perm[n_test:]
        tr = torch.as_tensor(np.sort(train_idx), device="cuda")
        model = sg.Surrogate(seed=0)
        iters = args.warmup if args.warmup is not None else 20000
        t0 = time.perf_counter()
        losses = sg.train(model, xs[tr], tg[tr], iters=iters, batch=256, lr=1e-3,
                          lr_final=5e-5, log=log, log_every=1000)
        train_s = time.perf_counter() - t0
        ev = sg.evaluate(model, xs, tg, test_idx, batch=64)
        log(f"[bench] held-out: {ev}")
        train_info = {"pairs": int(N), "train_iters": iters, "train_s": train_s,
                      "final_loss": losses[-1][1] if losses else None, "heldout_K": ev}
        if args.save_params:
            sg.save_params(model, args.save_params)
            log(f"[bench] parameters saved to {args.save_params}")
        del xs, tg
        torch.cuda.empty_cache()
    # MFMA forward throughput at batch 256 (the conv layers' roofline)
    x0 = torch.zeros((256, sg.WIN, sg.WIN, sg.CH), dtype=torch.bfloat16, device="cuda")
    for _ in range(2):
        model.forward(x0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        model.forward(x0)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 10
    del x0
    mfma_flop = 5 * 256 * sg.WIN * sg.WIN * sg.CH * 9 * sg.CH * 2   # the five MFMA convs
    act = sg.WIN * sg.WIN * sg.CH * 2                   # one bf16 activation tensor per window
    fwd_bytes = 256 * (5 * 2 * act + act + sg.WIN * sg.WIN * 4)

    rs = sg.RunSpec(steps=steps)
    mat, pr, dom, k, plan = sg.run_setup(rs)
    log("[bench] FP64 reference run")
    T64 = sg.fp64_reference(k, plan, steps)
    T_plain, w_plain = sg.plain_run(rs, log=log)
    T_fd, w_fd, c_fd, _ = sg.two_level_run(rs, None, log=log)
    T_net, w_net, c_net, diffs = sg.two_level_run(rs, model, log=log)
    sim = steps * rs.dt

    def err(T):
        e = (T - T64).abs()
        at = np.unravel_index(int(e.argmax()), e.shape)
        return {"max_K": float(e.max()), "mean_K": float(e.mean()),
                "at_cell_xy": [int(at[1]), int(at[0])]}

    out = {
        "metric": "sim-s/wall-s, learned two-level run (4096^2, 40k steps)",
        "value": sim / w_net, "unit": "sim-s/wall-s", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": w_net * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network) / f32 (FD)",
        "data": "synthetic",
        "config": {
            "workload": "4096^2 plate, dx 5 um, dt 1.2 us, 10-track raster P 300 W V 1 m/s, "
                        "40,000 steps; two-level run: far field on a 2x coarse grid (4 fine "
                        "steps per coarse step), 128^2 fine window (96 cells of trail) "
                        "advanced 16 steps per network call, seam ramp of 8 cells, FD recheck "
                        "every 8 calls, fallback above 2 K",
            "speedup_vs_plain": w_plain / w_net,
            "plain_fp32": {"sim_s_per_wall_s": sim / w_plain, "wall_s": w_plain,
                           "error_vs_fp64": err(T_plain)},
            "two_level_fd_window": {"sim_s_per_wall_s": sim / w_fd, "wall_s": w_fd,
                                    "speedup_vs_plain": w_plain / w_fd,
                                    "error_vs_fp64": err(T_fd), "counters": c_fd},
            "two_level_network": {"sim_s_per_wall_s": sim / w_net, "wall_s": w_net,
                                  "error_vs_fp64": err(T_net), "counters": c_net,
                                  "applied_rate": c_net["applied"] / max(c_net["calls"], 1),
                                  "fallback_rate": c_net["fallbacks"] / max(c_net["rechecks"], 1),
                                  "recheck_K": {"median": float(np.median(diffs)) if len(diffs)
                                                else None,
                                                "max": float(diffs.max()) if len(diffs) else None}},
            "peak_T_K_fp64": float(T64.max()),
            "training": train_info,
        },
        # 144 FLOP per byte (each layer reads its bf16 input once and writes its
        # output once) is below the MI355X ridge (2.5 PFLOP/s / 8 TB/s = 312), so
        # the forward is bounded by HBM; the MFMA fraction is reported beside it
        "roofline": {"bound": "hbm", "kernel": "k_conv_mfma x5 + k_out_fwd (batch 256 forward)",
                     "achieved": fwd_bytes / (fwd_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fwd_bytes / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes": fwd_bytes,
                     "mfma_TFLOPs": mfma_flop / (fwd_ms * 1e-3) / 1e12,
                     "mfma_frac": mfma_flop / (fwd_ms * 1e-3) / 1e12 / 2500.0,
                     "forward_ms_batch256": fwd_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def plate_ext_bench(args):
    """The extended physics (temperature-dependent IN718 properties, latent
    heat) on the `--n` x `--n` plate, dx 5 um, dt at the extended-physics
    stability limit, the beam parked at the plate centre with the plate's
    P and r0: the FP32 kernel (lpbf_launch_step_f32_ext, one step per launch,
    fused statistics) and the FP64 parity kernel (lpbf_launch_step_ex,
    phys = 1, block statistics), each over `--steps` steps after the
    warm-up.  value = FP32 cell-updates per second."""
    import numpy as np
    import torch

    import hedgehog_lpbf_amd as h
    lib = h.lib()
    mat, pr, dom, _, _, _ = plate_setup(args.n, args.n)
    nx, ny = dom.nx, dom.ny
    dt = h.dt_cfl_ext(mat, dom, 1.0)
    k = h.make_const(mat, pr, dom, dt)
    k.x_l, k.y_l, k.q_scale = 0.5 * dom.Lx, 0.5 * dom.Ly, 1.0
    steps = args.steps if args.steps is not None else 200
    warm = args.warmup if args.warmup is not None else 20
    plan = np.tile(np.array([[k.x_l, k.y_l, 1.0]]), (warm + steps, 1))
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream(dev).cuda_stream
    pdev = torch.from_numpy(plan).to(dev)
    ld = (nx + 3) // 4 * 4
    a = torch.empty((ny, ld), dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    lib.lpbf_launch_noise_f32(a.data_ptr(), nx, ny, ld, 0, 0.0, 0.5, st)
    work = torch.empty(lib.lpbf_f32_ext_workspace_bytes(nx, ny) // 8 + 8, dtype=torch.float64,
                       device=dev)
    stats = torch.empty((warm + steps) * 7, dtype=torch.float64, device=dev)
    bufs = [a, b]

    def f32(s0, n):
        for s in range(s0, s0 + n):
            lib.lpbf_launch_step_f32_ext(bufs[0].data_ptr(), bufs[1].data_ptr(), ld, k, 300.0, 0,
                                         300.0, mat.Tm, pdev.data_ptr(), s,
                                         stats.data_ptr() + 56 * s, work.data_ptr(), st)
            bufs.reverse()

    def timed(f, s0, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize(dev)
        e0.record()
        f(s0, n)
        e1.record()
        torch.cuda.synchronize(dev)
        h.check()
        return e0.elapsed_time(e1)

    f32(0, warm)
    ms32 = timed(f32, warm, steps)
    A = torch.empty((ny, nx), dtype=torch.float64, device=dev)
    B = torch.empty_like(A)
    lib.lpbf_launch_noise_f64(A.data_ptr(), nx * ny, 0, 300.0, 0.5, st)
    blk = torch.empty(lib.lpbf_step_ex_nblocks(nx, ny) * 7 + 7, dtype=torch.float64, device=dev)
    fb = [A, B]

    def f64(s0, n):
        for _ in range(n):
            lib.lpbf_launch_step_ex(fb[0].data_ptr(), fb[1].data_ptr(), k, 1, 0, 300.0, mat.Tm,
                                    blk.data_ptr(), st)
            fb.reverse()

    f64(0, max(2, warm // 4))
    n64 = max(10, steps // 4)
    ms64 = timed(f64, 0, n64)
    r32 = nx * ny * steps / (ms32 * 1e-3)
    r64 = nx * ny * n64 / (ms64 * 1e-3)
    print(json.dumps({
        "metric": "cell-updates/s", "value": r32, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": ms32 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nx}x{ny} plate, dx 5 um, extended physics (IN718 tables, "
                               "latent heat, kmult 1), dt at its stability limit, beam parked "
                               "at the centre, one step per launch with fused statistics",
                   "dt_s": dt, "fp64_cell_updates_per_s": r64, "fp64_ms_per_step": ms64 / n64,
                   "speedup_vs_fp64": r32 / r64},
        "roofline": {"bound": "valu", "kernel": "k_step_f32_ext",
                     "alg_bytes_per_step": 8 * nx * ny,
                     "alg_frac_hbm": 8 * nx * ny / (ms32 / steps * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": None}), flush=True)


def sg_kernels_bench(args):
    """The surrogate's kernels at batch 256 of 128 x 128 windows (32 channels,
    BF16 activations, FP32 accumulation), on synthetic windows and the initial
    parameters: the forward pass (five MFMA conv layers and the output layer)
    and the training pass (forward with z stored, loss, backward data and
    weight gradients), each timed over `--steps` repetitions.  value = forward
    windows per second; the roofline is the forward pass's algorithmic bytes
    (every layer reads its input once and writes its output once) over its
    time."""
    import torch

    from hedgehog_lpbf_amd import surrogate as sg
    B = 256
    reps = args.steps if args.steps is not None else 20
    torch.manual_seed(0)
    model = sg.Surrogate(seed=0)
    x0 = (torch.rand((B, sg.WIN, sg.WIN, sg.CH), device="cuda") - 0.5).to(torch.bfloat16)
    tg = torch.rand((B, sg.WIN, sg.WIN), device="cuda")

    def timed(f):
        for _ in range(args.warmup if args.warmup is not None else 2):
            f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd_ms = timed(lambda: model.forward(x0))
    train_ms = timed(lambda: model.loss_grad(x0, tg))
    act = sg.WIN * sg.WIN * sg.CH * 2
    fwd_bytes = B * (5 * 2 * act + act + sg.WIN * sg.WIN * 4)
    mfma_flop = 5 * B * sg.WIN * sg.WIN * sg.CH * 9 * sg.CH * 2
    print(json.dumps({
        "metric": "surrogate windows/s (forward, batch 256)", "value": B / (fwd_ms * 1e-3),
        "unit": "windows/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": fwd_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "surrogate network, batch 256 of 128x128 windows, 32 channels, "
                               "6 conv3x3 layers, initial parameters, random windows",
                   "forward_ms": fwd_ms, "training_pass_ms": train_ms},
        "roofline": {"bound": "hbm", "kernel": "k_conv_mfma x5 + k_out_fwd (batch 256 forward)",
                     "achieved": fwd_bytes / (fwd_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fwd_bytes / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes": fwd_bytes,
                     "mfma_TFLOPs": mfma_flop / (fwd_ms * 1e-3) / 1e12,
                     "mfma_frac": mfma_flop / (fwd_ms * 1e-3) / 1e12 / 2500.0,
                     "counter_summary": "profiles/k_conv_mfma_surrogate.md"},
        "cpu_baseline": None}), flush=True)


def copy_reference(args):
    """Plain device-to-device copy of an n x n FP32 plate, `steps` times."""
    import torch
    n = args.n
    a = torch.ones((n, n), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / max(args.steps, 1)
    nbytes = 2 * a.numel() * 4
    print(json.dumps({"metric": "copy bytes/s", "value": nbytes / (ms * 1e-3),
                      "unit": "B/s", "bytes_per_copy": nbytes, "read_bytes": nbytes // 2,
                      "ms_per_copy": ms, "steps": args.steps}), flush=True)


VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4   # wave64 VALU instructions/s: 1,024 SIMDs, 2.4 GHz, 4 cycles each


def plate_measure(args, precision, min_window_ms=200.0):
    """One plate line: autotune (unless --depth), warm-up, then timed windows
    of exactly `steps` steps each (stream synchronised before and after every
    window), repeated until `min_window_ms` of timed GPU work so that a short
    driver run is not a single millisecond; value from the median window.
    The step kernel's average launch duration comes from HIP events on its
    stream around every tenth full-depth launch of the timed windows."""
    import numpy as np
    import torch

    import hedgehog_lpbf_amd as h
    lib = h.lib()
    mat, pr, dom, dt, k, plan = plate_setup(args.n, args.n)
    nx, ny = dom.nx, dom.ny
    cells = nx * ny
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    st = stream.cuda_stream
    pdev = torch.from_numpy(plan).to(dev)
    stats = torch.empty(plan.shape[0] * 7 + 7, dtype=torch.float64, device=dev)
    tag = "plate8k_f32" if precision == 32 else "plate8k_f64"
    print(f"[bench] {tag} {nx}x{ny} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)
    depth, rows, tuned = args.depth, args.rows, None
    if depth == 0:
        bd, br, bms = C.c_int(), C.c_int(), C.c_double()
        t0 = time.perf_counter()
        if precision == 32:
            rc = lib.lpbf_f32_autotune(k, 300.0, 0, 300.0, mat.Tm, pdev.data_ptr(),
                                       plan.shape[0], 4, st, C.byref(bd), C.byref(br),
                                       C.byref(bms))
        else:
            rc = lib.lpbf_f64_blocked_autotune(k, 0, 0, 300.0, mat.Tm, pdev.data_ptr(),
                                               plan.shape[0], 4, st, C.byref(bd), C.byref(br),
                                               C.byref(bms))
        if rc != 0:
            h.check()
            raise RuntimeError(f"autotune failed: {rc}")
        depth, rows = bd.value, br.value
        tuned = {"depth": depth, "rows": rows, "ms_per_step": bms.value,
                 "tune_s": time.perf_counter() - t0}
        print(f"[bench] autotune: depth {depth} rows {rows} ({bms.value:.4f} ms/step)",
              file=sys.stderr, flush=True)
    if precision == 32:
        ld = (nx + 3) // 4 * 4
        a = torch.empty((ny, ld), dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        lib.lpbf_launch_noise_f32(a.data_ptr(), nx, ny, ld, 0, 0.0, 0.5, st)
        work = torch.empty(lib.lpbf_f32_workspace_bytes(nx, ny, depth, rows) // 8 + 8,
                           dtype=torch.float64, device=dev)
        bytes_per_update = 8

        def launch(s, d, x, y):
            lib.lpbf_launch_step_f32(x.data_ptr(), y.data_ptr(), ld, k, 300.0, 0, 300.0,
                                     mat.Tm, pdev.data_ptr(), s, d, rows, None, work.data_ptr(),
                                     st)

        def finalize(s, d):
            lib.lpbf_launch_f32_stats_finalize(work.data_ptr(), nx, ny, d, rows,
                                               stats.data_ptr() + 56 * s, st)
        kernel_name = "k_step_f32"
    else:
        # FP64 parity path with temporal blocking: the field is bitwise the
        # one `depth` single steps give
        a = torch.empty((ny, nx), dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        lib.lpbf_launch_noise_f64(a.data_ptr(), cells, 0, 300.0, 0.5, st)
        work = torch.empty(lib.lpbf_f64_blocked_workspace_bytes(nx, ny, depth, rows) // 8 + 8,
                           dtype=torch.float64, device=dev)
        bytes_per_update = 16

        def launch(s, d, x, y):
            lib.lpbf_launch_step_f64_blocked(x.data_ptr(), y.data_ptr(), k, 0, 0, 300.0, mat.Tm,
                                             pdev.data_ptr(), s, d, rows,
                                             stats.data_ptr() + 56 * s, work.data_ptr(), st)

        def finalize(s, d):
            pass                    # the blocked launch merges its statistics itself
        kernel_name = "k_step_f64_tb"
    bufs = [a, b]

    def run(s0, n, events=None, every=10):
        # events: pre-created (start, end, depth) slots recorded on the launch
        # stream around every `every`-th step launch (sparse, so that the
        # markers barely perturb the window they sit in)
        s, q, i = s0, 0, 0
        while s < s0 + n:
            d = min(depth, s0 + n - s)
            timed = events is not None and i % every == 0 and q < len(events)
            if timed:
                events[q][0].record(stream)
            launch(s, d, bufs[0], bufs[1])
            if timed:
                events[q][1].record(stream)
                events[q][2] = d
                q += 1
            finalize(s, d)
            bufs.reverse()
            s += d
            i += 1

    run(0, args.warmup)
    torch.cuda.synchronize(dev)
    h.check()
    # timed windows: events around every tenth launch give the step kernel's
    # duration in steady state (created before the window starts)
    windows, events = [], []
    s0 = args.warmup
    while True:
        if s0 + args.steps > plan.shape[0]:
            break
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        ev = [[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True), 0]
              for _ in range(-(-args.steps // depth) // 10 + 1)]
        events += ev
        torch.cuda.synchronize(dev)
        t_start.record(stream)
        run(s0, args.steps, ev)
        t_end.record(stream)
        torch.cuda.synchronize(dev)
        windows.append(t_start.elapsed_time(t_end))
        s0 += args.steps
        if sum(windows) >= min_window_ms or len(windows) >= 1000:
            break
    h.check()
    if not windows:
        raise SystemExit(f"--warmup + --steps exceed the {plan.shape[0]}-step scan plan")
    win_ms = float(np.median(windows))
    last = (h.StepStats * 1).from_buffer_copy(
        stats.cpu().numpy().tobytes()[56 * (s0 - 1): 56 * s0])[0]
    full = [e0.elapsed_time(e1) for e0, e1, d in events if d == depth]
    kern_ms = float(np.mean(full)) if full else float("nan")
    print(f"[bench] {len(windows)} windows of {args.steps} steps, median {win_ms:.3f} ms; "
          f"step-kernel avg {kern_ms:.4f} ms per launch of {depth} steps; peak T "
          f"{last.tmax:.1f} K, molten {last.molten}", file=sys.stderr, flush=True)

    # roofline from the committed counter summary of the same launch shape
    prof = counters_for(tag, depth=depth, rows=rows)
    alg_bytes = bytes_per_update * cells * depth     # per launch, counted per step
    compulsory = bytes_per_update * cells            # one read + one write per launch
    kern_s = kern_ms * 1e-3
    traffic = prof.get("bytes_per_launch") if prof else None
    basis = "counters (FETCH_SIZE + WRITE_SIZE, copy-calibrated)" if traffic else \
        "compulsory bytes (no counter summary for this launch shape)"
    hbm_bytes = traffic if traffic else compulsory
    achieved = hbm_bytes / kern_s / 1e9
    valu = prof.get("valu_insts_per_launch") if prof else None
    roof = {
        "bound": prof.get("bound", "hbm") if prof else "hbm",
        "kernel": kernel_name,
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "traffic_basis": basis,
        "valu_frac": (valu / (kern_s * VALU_ISSUE_PER_S)) if valu else None,
        "valu_insts_per_launch": valu,
        "alg_bytes_per_launch": alg_bytes,
        "alg_frac": alg_bytes / kern_s / (HBM_PEAK_GBS * 1e9),
        "kernel_ms_per_launch": kern_ms,
        "counter_summary": prof.get("summary") if prof else None,
    }
    return {
        "value": cells * args.steps / (win_ms * 1e-3),
        "ms_per_step": win_ms / args.steps,
        "windows": len(windows),
        "window_ms": {"min": float(min(windows)), "median": win_ms,
                      "max": float(max(windows))},
        "depth": depth, "rows": rows, "tuned": tuned, "roofline": roof,
        "k": k, "plan": plan, "nx": nx, "ny": ny,
        "peak_T_K": last.tmax,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps per window (default: 4000 for the plates, 5000 for "
                         "batch256, 20 repetitions for case0)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="plate8k_f32",
                    choices=["plate8k_f32", "plate8k_f64", "batch256", "case0", "ensemble",
                             "ensemble_f32", "run8k_f64", "run8k_f32", "surrogate",
                             "surrogate_kernels", "plate8k_f32_ext", "copy"],
                    help="plate8k_f32 (default): the FP32 plate, with the FP64 parity path "
                         "on the same plate in its `fp64` object; plate8k_f64: the FP64 "
                         "line alone; batch256: 256 uneven instances in one hipGraph launch; "
                         "case0: the reference's published case on the device-resident path; "
                         "ensemble: --runs manifest rows of that plate stepped together "
                         "(ensemble_f32: on the batched FP32 kernel); "
                         "run8k_f64 / run8k_f32: the device-resident run on the 8192^2 plate; "
                         "surrogate_kernels: the network's forward and training passes at "
                         "batch 256; plate8k_f32_ext: the extended physics on the plate, FP32 "
                         "beside FP64; "
                         "copy: a plain device copy of one n x n FP32 plate per step, the "
                         "reference for calibrating HBM byte counters")
    ap.add_argument("--depth", type=int, default=0,
                    help="plate workloads: steps per launch (1-4); 0 = pick depth and band "
                         "rows with the library's autotuner before the warm-up")
    ap.add_argument("--rows", type=int, default=0,
                    help="plate workloads: band height (0 = built-in)")
    ap.add_argument("--n", type=int, default=8192, help="plate side (cells)")
    ap.add_argument("--runs", type=int, default=64, help="ensemble: manifest rows")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-tiling", default="",
                    help="batch256: fixed depth:rows for size classes 0,1,2 (e.g. "
                         "4:16,3:48,3:48) instead of the autotuner")
    ap.add_argument("--no-fp64", action="store_true",
                    help="plate8k_f32: leave out the FP64 parity-path object")
    ap.add_argument("--data-steps", type=int, default=49168,
                    help="surrogate: FD steps per training run (64 runs; 49,168 gives about "
                         "196k window pairs)")
    ap.add_argument("--save-params", default=None,
                    help="surrogate: write the trained parameters (.npy) here")
    ap.add_argument("--load-params", default=None,
                    help="surrogate: skip training and load parameters (.npy)")
    ap.add_argument("--min-window-ms", type=float, default=200.0,
                    help="plate workloads: repeat the timed window of --steps steps until "
                         "this much GPU time is timed (value from the median window)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the benchmark "
              "runs on one GPU; the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import hedgehog_lpbf_amd as h

    h.lib()
    if args.workload == "batch256":
        return batch_bench(args)
    if args.workload == "case0":
        return case0_bench(args)
    if args.workload in ("run8k_f64", "run8k_f32"):
        return run8k_bench(args, 64 if args.workload == "run8k_f64" else 32)
    if args.workload in ("ensemble", "ensemble_f32"):
        return ensemble_bench(args, 32 if args.workload == "ensemble_f32" else 64)
    if args.workload == "surrogate_kernels":
        return sg_kernels_bench(args)
    if args.workload == "plate8k_f32_ext":
        return plate_ext_bench(args)
    if args.workload == "surrogate":
        return surrogate_bench(args)
    if args.steps is None:
        args.steps = 4000 if args.workload != "copy" else 10
    if args.warmup is None:
        args.warmup = 400 if args.workload != "copy" else 2
    if args.workload == "copy":
        return copy_reference(args)

    prec = 32 if args.workload == "plate8k_f32" else 64
    m = plate_measure(args, prec, args.min_window_ms)
    fp64 = None
    if prec == 32 and not args.no_fp64:
        args64 = argparse.Namespace(**vars(args))
        args64.depth, args64.rows = 0, 0
        m64 = plate_measure(args64, 64, args.min_window_ms)
        fp64 = {"metric": "cell-updates/s", "value": m64["value"], "unit": "cell-updates/s",
                "dtype": "f64", "steps": args.steps, "windows": m64["windows"],
                "ms_per_step": m64["ms_per_step"], "window_ms": m64["window_ms"],
                "config": {"workload": "the same plate on the FP64 parity path (temporally "
                                       "blocked, bitwise the single-step field)",
                           "temporal_depth": m64["depth"], "band_rows": m64["rows"],
                           "autotuned": m64["tuned"]},
                "roofline": m64["roofline"], "cpu_baseline": None}
    nx, ny = m["nx"], m["ny"]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(m["k"], m["plan"], nx, ny, fp32=prec == 32)
        if fp64 is not None:
            fp64["cpu_baseline"] = cpu_baseline(m["k"], m["plan"], nx, ny, fp32=False)
    out = {
        "metric": "cell-updates/s",
        "value": m["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{nx}x{ny} plate, dx 5 um, dt 1.2 us, 20-track raster P 250 W "
                        f"V 1.2 m/s, T0 300 K +/-0.5 K noise, "
                        f"{'FP32 theta field' if prec == 32 else 'FP64 field'}"
                        f", fused per-step stats",
            "nx": nx, "ny": ny, "temporal_depth": m["depth"], "band_rows": m["rows"],
            "autotuned": m["tuned"],
            "timed_windows": m["windows"], "window_ms": m["window_ms"],
        },
        "roofline": m["roofline"],
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        out["fp64"] = fp64
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
