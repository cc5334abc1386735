"""FP64 parity of the HIP kernels against the CPU oracle (through the C ABI).

Bar: max |dT| <= 1e-9 K against the oracle for every field (the reference
north-star bound for the FP64 GPU path); integer/index results (argmax,
molten rows, partial records) bit-exact.  The oracle itself is pinned on
reference/case0 (tests/test_oracle_golden.py).
"""
import csv
import os

import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from tests.devutil import (dev, dev_bytes, kconst, partials_np, stats_np, stream,
                           sync_check)

pytestmark = pytest.mark.gpu
TOL = 1e-9
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _rand_field(ny, nx, seed, lo=300.0, hi=2600.0):
    return lo + (hi - lo) * np.random.default_rng(seed).random((ny, nx))


def _case(ora, nx=240, ny=120, case="0"):
    s = ora.defaults(case)
    s.dom.nx, s.dom.ny = nx, ny
    return s


@pytest.mark.parametrize("nx,ny", [(240, 120), (2, 2), (37, 5), (129, 67), (300, 33)])
@pytest.mark.parametrize("phys", [0, 1])
def test_step_single_parity(ora, nx, ny, phys):
    print(f"\n[case] step nx={nx} ny={ny} phys={phys}", flush=True)
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s) if phys == 0 else ora.dt_cfl_ext(s, 1.8))
    k.kmult = 1.8 if phys else 1.0
    k.x_l, k.y_l = 0.37 * s.dom.Lx, 0.41 * s.dom.Ly
    T = _rand_field(ny, nx, nx * 7 + ny)
    ref = ora.step(T, k, phys)
    a, b = dev(T), dev(np.zeros_like(T))
    kk = kconst(k)
    h.lib().lpbf_launch_step(a.data_ptr(), b.data_ptr(), kk, phys, stream())
    sync_check()
    got = b.cpu().numpy()
    assert np.abs(got - ref).max() <= TOL


@pytest.mark.parametrize("phys", [0, 1])
def test_step_bitwise_without_exp(ora, phys):
    """With the beam far outside the plate no cell calls exp, the one libm
    function the device and glibc may round differently, so the step is
    bit-identical to the oracle.  Pins the kernel's divisions (reciprocal
    plus one remainder correction), the property tables over their whole
    range (300..2600 K, the mushy zone included) and the row skip of the
    source."""
    print(f"\n[case] step bitwise without exp phys={phys}", flush=True)
    nx, ny = 300, 133
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s) if phys == 0 else ora.dt_cfl_ext(s, 1.8))
    k.kmult = 1.8 if phys else 1.0
    k.x_l, k.y_l = 0.5 * s.dom.Lx, 40.0 * s.dom.Ly
    T = _rand_field(ny, nx, 1234 + phys)
    ref = ora.step(T, k, phys)
    a, b = dev(T), dev(np.zeros_like(T))
    h.lib().lpbf_launch_step(a.data_ptr(), b.data_ptr(), kconst(k), phys, stream())
    sync_check()
    got = b.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), \
        int((got != ref).sum())


def test_case0_full_run_fused_path(ora):
    """The reference GpuStepTask fused mode (gpu_step_task.hpp:149-186):
    step_reduce + device finalize + hot-column extraction every step, run
    for the full case0 (1427 steps), against the oracle and the golden rows."""
    print("\n[case] case0 full run, fused step_reduce path", flush=True)
    s = ora.defaults("0")
    dt = ora.dt_cfl(s)
    n = int(1.5e-3 / dt)
    k = ora.make_const(s, dt)
    nx, ny = s.dom.nx, s.dom.ny
    T_ref, pk, W, Ws, summ = ora.run(s, dt, n)
    ticks = ora.build_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n, 200)
    lib = h.lib()
    nf = ((nx + 15) // 16) * ((ny + 15) // 16)
    a = dev(np.full((ny, nx), s.proc.Tamb))
    b = dev(np.zeros((ny, nx)))
    part = dev_bytes(nf * 24)
    glob = dev_bytes(n * 24)
    cols = dev(np.zeros((n, ny)))
    kk = kconst(k)
    st = stream()
    for i in range(n):
        kk.x_l, kk.y_l = ticks[i].x_l, ticks[i].y_l
        lib.lpbf_launch_step_reduce(a.data_ptr(), b.data_ptr(), kk, 0, 1609.0,
                                    part.data_ptr(), st)
        a, b = b, a
        lib.lpbf_launch_reduce_finalize(part.data_ptr(), nf, nx, ny, glob.data_ptr() + 24 * i, st)
        lib.lpbf_launch_extract_col(a.data_ptr(), nx, ny, glob.data_ptr() + 24 * i,
                                    cols.data_ptr() + 8 * ny * i, st)
    sync_check()
    final = a.cpu().numpy()
    assert np.abs(final - T_ref).max() <= TOL
    g = partials_np(glob, n)
    C = cols.cpu().numpy()
    dy = s.dom.Ly / (ny - 1)
    for i in range(n):
        assert abs(g[i].tmax - pk[i]) <= TOL
        Wi = 0.0 if g[i].ymax < 0 else (g[i].ymax - g[i].ymin + 1) * dy * 1e6
        assert Wi == W[i]
        wsub = ora.wsub_from_col(C[i], dy, 1609.0) if g[i].tmax > 1609.0 else 0.0
        assert abs(wsub - Ws[i]) <= 1e-6
    with open(os.path.join(GOLD, "case0_metrics.csv")) as f:
        for r in csv.DictReader(f):
            i = int(r["step"])
            assert "%g" % g[i].tmax == r["peakT_K"]


def test_reduce_grid_stride_records_bitwise(ora):
    print("\n[case] reduce grid-stride records", flush=True)
    nx, ny = 240, 120
    T = _rand_field(ny, nx, 3)
    T[50, 7] = T[90, 100] = T.max() + 10.0
    for nblocks in (1, 7, 113, 256):
        parts = ora.reduce_partials(T, 1609.0, nblocks)
        out = dev_bytes(nblocks * 24)
        h.lib().lpbf_launch_reduce(dev(T).data_ptr(), nx, ny, 1609.0, out.data_ptr(), nblocks,
                                   stream())
        sync_check()
        got = partials_np(out, nblocks)
        for b in range(nblocks):
            assert (got[b].tmax, got[b].arg, got[b].ymin, got[b].ymax) == \
                   (parts[b].tmax, parts[b].arg, parts[b].ymin, parts[b].ymax)


@pytest.mark.parametrize("nx,ny", [(16, 16), (17, 3), (240, 120), (700, 41)])
def test_step_reduce_fills_reference_partial_count(ora, nx, ny):
    print(f"\n[case] step_reduce nx={nx} ny={ny}", flush=True)
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 0.5 * s.dom.Lx, 0.5 * s.dom.Ly
    T = _rand_field(ny, nx, nx + ny, 300.0, 2000.0)
    ref = ora.step(T, k)
    nf = ((nx + 15) // 16) * ((ny + 15) // 16)
    part = dev(np.full(nf * 3 + 3, np.nan))   # poison: every record must be written
    a, b = dev(T), dev(np.zeros_like(T))
    out = dev_bytes(24)
    st = stream()
    h.lib().lpbf_launch_step_reduce(a.data_ptr(), b.data_ptr(), kconst(k), 0, 1609.0,
                                    part.data_ptr(), st)
    h.lib().lpbf_launch_reduce_finalize(part.data_ptr(), nf, nx, ny, out.data_ptr(), st)
    sync_check()
    assert np.abs(b.cpu().numpy() - ref).max() <= TOL
    recs = partials_np(part, nf)
    assert all(not np.isnan(r.tmax) for r in recs)
    g = partials_np(out, 1)[0]
    want = ora.reduce_finalize(ora.reduce_partials(b.cpu().numpy(), 1609.0, 4), nx, ny)
    assert (g.tmax, g.arg, g.ymin, g.ymax) == (want.tmax, want.arg, want.ymin, want.ymax)


def test_step_tile_matches_oracle_tiles(ora):
    print("\n[case] tile step", flush=True)
    s = _case(ora, 96, 40)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 0.3 * s.dom.Lx, 0.6 * s.dom.Ly
    T = _rand_field(40, 96, 9)
    y0s, y1s = ora.tile_bounds(40, 6)
    for phys in (0, 1):
        for y0, y1 in zip(y0s, y1s):
            tile = ora.make_tile(T, y0, y1)
            ref = ora.step_tile(tile, y1 - y0 + 1, y0, k, phys)
            a, b = dev(tile), dev(np.zeros_like(tile))
            h.lib().lpbf_launch_step_tile(a.data_ptr(), b.data_ptr(), y1 - y0 + 1, y0,
                                          kconst(k), phys, stream())
            sync_check()
            got = b.cpu().numpy()
            assert np.abs(got - ref).max() <= TOL
            assert np.array_equal(got[0], tile[0]) and np.array_equal(got[-1], tile[-1])


@pytest.mark.parametrize("bc", [0, 1])
def test_step_ex_stats(ora, bc):
    print(f"\n[case] step_ex bc={bc}", flush=True)
    nx, ny = 333, 97
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 0.5 * s.dom.Lx, 0.5 * s.dom.Ly
    T = _rand_field(ny, nx, 21, 300.0, 1900.0)
    ref, rst = ora.step_ex(T, k, 0, bc, 300.0, 1609.0)
    nb = h.lib().lpbf_step_ex_nblocks(nx, ny)
    blk = dev_bytes(nb * 56)
    out = dev_bytes(56)
    a, b = dev(T), dev(np.zeros_like(T))
    st = stream()
    h.lib().lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kconst(k), 0, bc, 300.0,
                                1609.0, blk.data_ptr(), st)
    h.lib().lpbf_launch_stats_finalize(blk.data_ptr(), nb, nx, ny, out.data_ptr(), st)
    sync_check()
    assert np.abs(b.cpu().numpy() - ref).max() <= TOL
    g = stats_np(out, 1)[0]
    assert abs(g.tmax - rst.tmax) <= TOL and g.arg == rst.arg
    assert (g.ymin, g.ymax, g.molten) == (rst.ymin, rst.ymax, rst.molten)
    for f in ("sum_dT", "sum_src", "sum_loss"):
        assert getattr(g, f) == pytest.approx(getattr(rst, f), rel=1e-11, abs=1e-9)


def test_config1_dirichlet_512_full_run(ora):
    """512x512, dx = 10 um, k=7 rho=4430 cp=526, h = 50 um, dt = 5 us,
    Dirichlet 300 K, P=200 W, eta=0.35, r0=50 um, V=1 m/s, 2000 steps."""
    print("\n[case] 512^2 Dirichlet FP64 2000 steps", flush=True)
    s = ora.defaults(None)
    s.mat.k, s.mat.rho, s.mat.cp = 7.0, 4430.0, 526.0
    s.proc.P, s.proc.V, s.proc.r0, s.proc.h = 200.0, 1.0, 50e-6, 50e-6
    s.proc.Tamb = s.proc.Tsub = 300.0
    s.dom.nx = s.dom.ny = 512
    s.dom.Lx = s.dom.Ly = 511 * 1e-5
    dt, n = 5e-6, 2000
    k = ora.make_const(s, dt)
    plan = ora.track_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n)
    T0 = np.full((512, 512), 300.0)
    Tref, rstats = ora.run_plan(k, plan, T0, 0, 1, 300.0, 1609.0)
    nb = h.lib().lpbf_step_ex_nblocks(512, 512)
    blk = dev_bytes(nb * 56)
    sts = dev_bytes(n * 56)
    a, b = dev(T0), dev(T0)
    kk = kconst(k)
    st = stream()
    for i in range(n):
        kk.x_l, kk.y_l, kk.q_scale = plan[i]
        h.lib().lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kk, 0, 1, 300.0, 1609.0,
                                    blk.data_ptr(), st)
        h.lib().lpbf_launch_stats_finalize(blk.data_ptr(), nb, 512, 512,
                                           sts.data_ptr() + 56 * i, st)
        a, b = b, a
    sync_check()
    got = a.cpu().numpy()
    assert np.abs(got - Tref).max() <= TOL
    gs = stats_np(sts, n)
    for i in range(0, n, 97):
        assert abs(gs[i].tmax - rstats[i].tmax) <= TOL and gs[i].arg == rstats[i].arg
        assert gs[i].molten == rstats[i].molten
        assert gs[i].sum_src == pytest.approx(rstats[i].sum_src, rel=1e-11)


def test_noise_f64_bitwise(ora):
    print("\n[case] noise f64", flush=True)
    n = 100003
    t = dev(np.zeros(n))
    h.lib().lpbf_launch_noise_f64(t.data_ptr(), n, 42, 300.0, 0.5, stream())
    sync_check()
    assert np.array_equal(t.cpu().numpy(), ora.noise_field(42, 300.0, 0.5, (n,)))


def test_invalid_argument_reported_without_launch():
    h.lib().lpbf_get_last_error()
    k = h.StepConst()
    k.nx, k.ny = 1, 5
    h.lib().lpbf_launch_step(16, 32, k, 0, None)
    with pytest.raises(h.LpbfError):
        h.check()


@pytest.mark.parametrize("bc", [0, 1])
def test_step_ex_batch_matches_step_ex(ora, bc):
    """The counter-free batch launches (step j plus an extra workgroup that
    finishes step j-1, then a flush) give the field and hot columns of
    lpbf_launch_step_ex + lpbf_launch_stats_finalize + lpbf_launch_extract_col
    bit for bit, and the same per-step max, index, molten rows and count; the
    FP64 energy sums agree to rounding (the batch form uses shorter bands, so
    its block records split the sums differently)."""
    print(f"\n[case] step_ex_batch bc={bc}", flush=True)
    nx, ny, n = 240, 120, 6
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s))
    T = _rand_field(ny, nx, 77, 300.0, 1900.0)
    plan = np.zeros((n, 3))
    plan[:, 0] = (0.3 + 0.02 * np.arange(n)) * s.dom.Lx
    plan[:, 1] = 0.5 * s.dom.Ly
    plan[:, 2] = 1.0
    cruise = np.array([1, 0, 1, 1, 0, 1], dtype=np.uint8)
    st = stream()
    lib = h.lib()
    # reference sequence
    nb = lib.lpbf_step_ex_nblocks(nx, ny)
    blk = dev_bytes(nb * 56)
    ring_r = dev_bytes(56 * n)
    cols_r = dev(np.zeros((n, ny)))
    a, b = dev(T), dev(np.zeros_like(T))
    kk = kconst(k)
    for j in range(n):
        kk.x_l, kk.y_l, kk.q_scale = plan[j]
        lib.lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kk, 0, bc, 300.0, 1609.0,
                                blk.data_ptr(), st)
        a, b = b, a
        lib.lpbf_launch_stats_finalize(blk.data_ptr(), nb, nx, ny, ring_r.data_ptr() + 56 * j, st)
        if cruise[j]:
            lib.lpbf_launch_extract_col(a.data_ptr(), nx, ny, ring_r.data_ptr() + 56 * j,
                                        cols_r.data_ptr() + 8 * ny * j, st)
    field_r = a
    # batch sequence
    nbb = lib.lpbf_step_ex_batch_nblocks(nx, ny)
    blk2 = dev_bytes(2 * nbb * 56)
    ring = dev_bytes(56 * n)
    cols = dev(np.zeros((n, ny)))
    ticks = dev(plan)
    cr = dev(cruise)
    a, b = dev(T), dev(np.zeros_like(T))
    for j in range(n):
        lib.lpbf_launch_step_ex_batch(a.data_ptr(), b.data_ptr(), kconst(k), 0, bc, 300.0, 1609.0,
                                      ticks.data_ptr(), j, cr.data_ptr(), blk2.data_ptr(),
                                      ring.data_ptr(), cols.data_ptr(), st)
        a, b = b, a
    lib.lpbf_launch_step_ex_batch_flush(a.data_ptr(), nx, ny, n, cr.data_ptr(), blk2.data_ptr(),
                                        ring.data_ptr(), cols.data_ptr(), st)
    sync_check()
    assert np.array_equal(a.cpu().numpy().view(np.uint64), field_r.cpu().numpy().view(np.uint64))
    for g, r in zip(stats_np(ring, n), stats_np(ring_r, n)):
        assert (g.tmax, g.arg, g.ymin, g.ymax, g.molten) == (r.tmax, r.arg, r.ymin, r.ymax,
                                                             r.molten)
        for f in ("sum_dT", "sum_src", "sum_loss"):
            assert getattr(g, f) == pytest.approx(getattr(r, f), rel=1e-12, abs=1e-12)
    got_c, ref_c = cols.cpu().numpy(), cols_r.cpu().numpy()
    for j in range(n):
        if cruise[j]:
            assert np.array_equal(got_c[j], ref_c[j]), j


def test_step_ex_plan_ring_matches_step_ex(ora):
    """The ticket-merged plan-driven step (lpbf_launch_step_ex_plan): six
    launches advance the device step counter, fill the ring and copy the hot
    columns on cruise steps; field and columns equal the step_ex path bit for
    bit, the ring's max, index, rows and molten count exactly, its sums to
    rounding; the counter ends at six and the ticket back at zero."""
    print("\n[case] step_ex_plan ring", flush=True)
    nx, ny, n = 240, 120, 6
    s = _case(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s))
    T = _rand_field(ny, nx, 78, 300.0, 1900.0)
    plan = np.zeros((n, 3))
    plan[:, 0] = (0.3 + 0.02 * np.arange(n)) * s.dom.Lx
    plan[:, 1] = 0.5 * s.dom.Ly
    plan[:, 2] = 1.0
    cruise = np.array([0, 1, 1, 0, 1, 1], dtype=np.uint8)
    st = stream()
    lib = h.lib()
    nb = lib.lpbf_step_ex_nblocks(nx, ny)
    blk = dev_bytes(nb * 56)
    ring_r = dev_bytes(56 * n)
    cols_r = dev(np.zeros((n, ny)))
    a, b = dev(T), dev(np.zeros_like(T))
    kk = kconst(k)
    for j in range(n):
        kk.x_l, kk.y_l, kk.q_scale = plan[j]
        lib.lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kk, 0, 0, 300.0, 1609.0,
                                blk.data_ptr(), st)
        a, b = b, a
        lib.lpbf_launch_stats_finalize(blk.data_ptr(), nb, nx, ny, ring_r.data_ptr() + 56 * j, st)
        if cruise[j]:
            lib.lpbf_launch_extract_col(a.data_ptr(), nx, ny, ring_r.data_ptr() + 56 * j,
                                        cols_r.data_ptr() + 8 * ny * j, st)
    field_r = a
    npb = lib.lpbf_step_ex_plan_nblocks(nx, ny)
    blk2 = dev_bytes(npb * 56)
    ring = dev_bytes(56 * n)
    cols = dev(np.zeros((n, ny)))
    counters = dev(np.zeros(2, dtype=np.int32))      # step counter, ticket
    pd, cr = dev(plan), dev(cruise)
    a, b = dev(T), dev(np.zeros_like(T))
    for j in range(n):
        lib.lpbf_launch_step_ex_plan(a.data_ptr(), b.data_ptr(), kconst(k), 0, 0, 300.0, 1609.0,
                                     pd.data_ptr(), counters.data_ptr(), cr.data_ptr(),
                                     blk2.data_ptr(), counters.data_ptr() + 4, ring.data_ptr(), n,
                                     cols.data_ptr(), st)
        a, b = b, a
    sync_check()
    assert counters.cpu().numpy().tolist() == [n, 0]
    assert np.array_equal(a.cpu().numpy().view(np.uint64), field_r.cpu().numpy().view(np.uint64))
    for g, r in zip(stats_np(ring, n), stats_np(ring_r, n)):
        assert (g.tmax, g.arg, g.ymin, g.ymax, g.molten) == (r.tmax, r.arg, r.ymin, r.ymax,
                                                             r.molten)
        for f in ("sum_dT", "sum_src", "sum_loss"):
            assert getattr(g, f) == pytest.approx(getattr(r, f), rel=1e-12, abs=1e-12)
    got_c, ref_c = cols.cpu().numpy(), cols_r.cpu().numpy()
    for j in range(n):
        if cruise[j]:
            assert np.array_equal(got_c[j], ref_c[j]), j


def test_autotuners_pick_valid_tilings(ora):
    """lpbf_f32_autotune and lpbf_f64_blocked_autotune on a small plate:
    they return a depth in range and a band height the launchers accept, and
    report the error state clean."""
    import ctypes as C
    print("\n[case] autotuners", flush=True)
    s = _case(ora, 300, 200)
    k = kconst(ora.make_const(s, ora.dt_cfl(s)))
    plan = np.zeros((8, 3))
    plan[:, 0], plan[:, 1], plan[:, 2] = 0.4 * s.dom.Lx, 0.5 * s.dom.Ly, 1.0
    pd = dev(plan)
    lib = h.lib()
    d, r, ms = C.c_int(), C.c_int(), C.c_double()
    assert lib.lpbf_f32_autotune(k, 300.0, 0, 300.0, 1609.0, pd.data_ptr(), 8, 4, stream(),
                                 C.byref(d), C.byref(r), C.byref(ms)) == 0
    assert 1 <= d.value <= 4 and r.value >= 0 and ms.value > 0
    assert lib.lpbf_f32_workspace_bytes(300, 200, d.value, r.value) > 0
    assert lib.lpbf_f64_blocked_autotune(k, 0, 0, 300.0, 1609.0, pd.data_ptr(), 8, 4, stream(),
                                         C.byref(d), C.byref(r), C.byref(ms)) == 0
    assert 1 <= d.value <= 4 and r.value >= 0 and ms.value > 0
    assert lib.lpbf_f64_blocked_workspace_bytes(300, 200, d.value, r.value) > 0
    sync_check()
