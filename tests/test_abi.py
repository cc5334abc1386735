"""CPU-side checks of the C ABI: the library loads, exports every symbol the
public headers declare, rejects invalid arguments without launching, and
its host arithmetic (make_const, dt rules, laser plans) equals the oracle's
bit for bit.  No GPU is touched."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import hedgehog_lpbf_amd as h

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(hdr).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(lpbf_[A-Za-z0-9_]+)\s*\(", txt))
    return names


def test_run_options_layout():
    assert C.sizeof(h.RunOptions) == 64 and C.sizeof(h.StepRecord) == 80


def test_library_exports_every_declared_symbol():
    lib = h.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(h._lib.SIGNATURES) == names


def test_invalid_arguments_are_recorded():
    lib = h.lib()
    lib.lpbf_get_last_error()
    k = h.StepConst()
    k.nx, k.ny = 1, 4
    lib.lpbf_launch_step(16, 32, k, 0, None)
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
    assert lib.lpbf_get_last_error() == 0
    k.nx = 64
    lib.lpbf_launch_step(16, 32, k, 7, None)        # bad phys
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
    lib.lpbf_launch_step_f32(16, 16, 64, k, 300.0, 0, 0.0, 1609.0, 16, 0, 1, 0,
                             None, 16, None)                      # in == out
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
    lib.lpbf_launch_step_f32(16, 48, 62, k, 300.0, 0, 0.0, 1609.0, 16, 0, 5, 0,
                             None, 16, None)                      # depth 5, ld < nx
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
    assert "invalid" in lib.lpbf_error_string(1000).decode()
    with pytest.raises(h.LpbfError):
        lib.lpbf_launch_reduce(16, 0, 4, 1609.0, 32, 1, None)
        h.check()


def test_run_create_rejects_bad_options_before_touching_a_device():
    k = h.StepConst()
    k.nx, k.ny, k.dt = 64, 32, 1e-6
    plan = np.zeros((4, 3))
    for kw in (dict(precision=16), dict(bc=5), dict(precision=32, depth=9), dict(rows=-1)):
        with pytest.raises(h.LpbfError):
            h.DeviceRun(k, Tm=1609.0, Lx=1e-3, Ly=1e-3, Tamb=300.0, plan=plan, **kw)
    assert h.lib().lpbf_get_last_error() == 0


def test_ensemble_create_rejects_bad_runs_before_touching_a_device():
    def member(nx=64, ny=32, n=4):
        k = h.StepConst()
        k.nx, k.ny, k.dt = nx, ny, 1e-6
        return dict(k=k, plan=np.zeros((n, 3)))
    for runs, kw in (([], {}),                                  # no runs
                     ([member()], dict(phys=3)),                 # unknown physics
                     ([member()], dict(bc=5)),                   # unknown boundary
                     ([member()], dict(batch=0)),
                     ([member()], dict(precision=32, phys=1)),   # FP32: baseline only
                     ([member()], dict(precision=32, depth=5)),
                     ([member()], dict(precision=32, depth=4, rows=6)),   # rows < 2*depth
                     ([member(), member(nx=48)], {})):           # grids differ
        with pytest.raises((h.LpbfError, ValueError)):
            h.Ensemble(runs, Tm=1609.0, Lx=1e-3, Ly=1e-3, **kw)
    h.lib().lpbf_get_last_error()
    # a precision other than 64 or 32 is refused by the C entry point itself
    lib, k = h.lib(), (h.StepConst * 1)(member()["k"])
    opt = h.RunOptions(16, 0, 300.0, 64, 4, 0, -1.0, 0, 1, 0, 0, 0)
    n = (C.c_int * 1)(4)
    plan = np.zeros((4, 3))
    out = C.c_void_p()
    assert lib.lpbf_ensemble_create(k, 1, 0, 1609.0, 1e-3, 1e-3, None, plan.ctypes.data, None,
                                    None, n, C.byref(opt), C.byref(out)) == 1000
    assert not out.value
    assert lib.lpbf_get_last_error() == 1000
    assert lib.lpbf_ensemble_advance(None, 1) == -1000
    lib.lpbf_get_last_error()


def test_batch_create_rejects_bad_instances_before_touching_a_device():
    lib = h.lib()
    arr = (h._lib.F32Instance * 2)()
    for a in arr:
        a.k.nx, a.k.ny, a.k.dt = 64, 32, 1e-6
        a.theta0, a.theta1, a.ld, a.d_plan, a.n_steps = 256, 512, 64, 1024, 10
    out = C.c_void_p()
    assert lib.lpbf_f32_batch_create(arr, 2, 7, C.byref(out)) == 1000          # bad bc
    arr[1].ld = 62                                                           # ld < nx
    assert lib.lpbf_f32_batch_create(arr, 2, 0, C.byref(out)) == 1000
    arr[1].ld, arr[1].theta1 = 64, 256                                       # aliasing
    assert lib.lpbf_f32_batch_create(arr, 2, 0, C.byref(out)) == 1000
    assert lib.lpbf_f32_batch_create(arr, 0, 0, C.byref(out)) == 1000
    assert lib.lpbf_get_last_error() == 1000
    assert lib.lpbf_f32_batch_enqueue(None, 0, 1, None) == 1000
    assert lib.lpbf_f32_batch_enqueue_launch(None, 0, 1, None) == 1000
    assert lib.lpbf_launch_hot_cols_f32(None, 1, 0, 0, 0, 0, None) == 1000
    assert lib.lpbf_launch_hot_cols_f32(1024, 1, 2, 0, 0, 0, None) == 1000     # q not 0/1
    lib.lpbf_get_last_error()


def test_make_const_and_dt_equal_oracle(ora):
    for case in ("0", "1.1", "3.2"):
        s = ora.defaults(case)
        mat = h.Material()
        pr = h.apply_case(case, h.Process())
        dom = h.Domain()
        assert h.dt_cfl(mat, dom) == ora.dt_cfl(s)
        assert h.dt_cfl_ext(mat, dom, 2.0) == ora.dt_cfl_ext(s, 2.0)
        dt = ora.dt_cfl(s)
        a, b = h.make_const(mat, pr, dom, dt), ora.make_const(s, dt)
        assert bytes(a) == bytes(b)


def test_plans_equal_oracle(ora):
    s = ora.defaults("0")
    dt = ora.dt_cfl(s)
    n = 1427
    plan, x_csv, wrapped = h.single_track_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n)
    assert np.array_equal(plan, ora.track_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n))
    ticks = ora.build_plan(s.proc.V, dt, s.dom.Lx, s.dom.Ly, n, 200)
    assert np.array_equal(x_csv, [t.x_laser_csv for t in ticks])
    assert np.array_equal(wrapped, [bool(t.wrapped) for t in ticks])
    r = h.centred_raster(40.96e-3, 40.96e-3, 1.2, 1.2e-6, 50000, 20, 80e-6)
    got = h.raster_plan(r["x0"], r["y0"], r["length"], r["hatch"], r["n_tracks"], r["V"],
                        1.2e-6, 50000)
    want = ora.raster_plan(ora.OraRaster(**r), 1.2e-6, 50000)
    assert np.array_equal(got, want)
    assert got[:, 2].sum() > 0.9 * 50000 and got[-1, 2] in (0.0, 1.0)


def _tb_records(nx, ny, d, rows):
    """Block records one blocked FP64 launch of depth d writes: d per
    workgroup of four strip-bands (strips of 128 loaded columns overlapping by
    2*ceil(d/2) per side; built-in band height: about four rounds of 3 * 1024
    wavefronts, at least 8*d rows)."""
    out = 128 - 4 * ((d + 1) // 2)
    nstrips = -(-nx // out)
    if rows == 0:
        bands = -(-(4 * 3 * 1024) // nstrips)
        rows = min(max(-(-ny // bands), 8 * d), ny)
    tasks = nstrips * -(-ny // rows)
    return d * -(-tasks // 4)


@pytest.mark.parametrize("nx,ny", [(240, 120), (512, 512), (2048, 1024), (8192, 8192),
                                   (1001, 77), (130, 3000), (64, 9)])
@pytest.mark.parametrize("rows", [0, 16, 33])
def test_f64_blocked_workspace_covers_shorter_launches(nx, ny, rows):
    """A device-resident run sizes the blocked kernel's workspace for its
    deepest launch and then also issues shorter ones (at cruise and dump
    steps): the size for depth D must hold the records of every depth d <= D.
    With the built-in band height a shorter depth can need more records."""
    lib = h.lib()
    sz = C.sizeof(h.StepStats)
    for D in range(1, 5):
        ws = lib.lpbf_f64_blocked_workspace_bytes(nx, ny, D, rows)
        for d in range(1, D + 1):
            assert ws >= sz * _tb_records(nx, ny, d, rows), (D, d)
    assert lib.lpbf_f64_blocked_workspace_bytes(nx, ny, 1, rows) == sz * _tb_records(nx, ny, 1, rows)
    # the cases where depth-D sizing alone was short
    if (nx, ny, rows) == (240, 120, 0):
        assert _tb_records(240, 120, 3, 0) > _tb_records(240, 120, 4, 0)


def test_noalloc_step_refuses_a_launch_beyond_its_scratch():
    """lpbf_launch_step_f32_noalloc checks its workspace need on the host and
    refuses, before any device work, a launch whose records exceed the fixed
    scratch (here 2-row bands of a 2^15-wide, 2^14-tall plate at depth 4)."""
    lib = h.lib()
    lib.lpbf_get_last_error()
    k = h.StepConst()
    k.nx, k.ny = 1 << 15, 1 << 14
    lib.lpbf_launch_step_f32_noalloc(16, 1 << 20, k.nx, k, 300.0, 0, 300.0, 1609.0, 16, 0, 4, 8,
                                     None, None)
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_SCRATCH_LIMIT
    assert "scratch" in lib.lpbf_error_string(h._lib.LPBF_ERR_SCRATCH_LIMIT).decode()
    lib.lpbf_launch_step_f32_noalloc(16, 16, 64, k, 300.0, 0, 300.0, 1609.0, 16, 0, 1, 0,
                                     None, None)                  # in == out
    assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
