"""FP32 extended physics (lpbf_launch_step_f32_ext, and DeviceRun with
precision=32, phys=1) against the FP64 oracle of the same physics
(oracle/lpbf_oracle.c, the restatement of step_task.hpp:132-187).

The FP32 path stores theta = T - T_ref, interpolates the IN718 table with
branch-free ramps and uses the hardware reciprocal, so its results are not
bitwise the FP64 ones.  The bounds below are its stated error: per step of a
field that spans the table and the mushy interval, and over whole runs of
the case0 plate with kmult 1 and 2.5, with mirror and Dirichlet edges."""
import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from tests.devutil import dev, dev_bytes, kconst, plan_dev, stats_np, stream, sync_check
from tests.test_gpu_f32 import _init, _setup

pytestmark = pytest.mark.gpu


def run_f32_ext(k, plan, T0, n, bc=0, T_bc=300.0, Tm=1609.0, T_ref=300.0):
    lib = h.lib()
    nx, ny = k.nx, k.ny
    ld = (nx + 3) // 4 * 4
    th = np.zeros((ny, ld), dtype=np.float32)
    th[:, :nx] = (T0 - T_ref).astype(np.float32)
    a = dev(th)
    b = a.clone()
    work = dev_bytes(lib.lpbf_f32_ext_workspace_bytes(nx, ny))
    stats = dev_bytes(max(n, 1) * 56)
    pdev, kk, st = plan_dev(plan), kconst(k), stream()
    for s in range(n):
        lib.lpbf_launch_step_f32_ext(a.data_ptr(), b.data_ptr(), ld, kk, T_ref, bc, T_bc, Tm,
                                     pdev.data_ptr(), s, stats.data_ptr() + 56 * s,
                                     work.data_ptr(), st)
        a, b = b, a
    sync_check()
    T = a[:, :nx].cpu().numpy().astype(np.float64) + T_ref
    return T, stats_np(stats, n)


@pytest.mark.parametrize("nx,ny,kmult,bc", [(240, 120, 1.0, 0), (240, 120, 2.5, 0),
                                            (333, 97, 1.0, 1), (129, 64, 2.5, 1)])
def test_f32_ext_steps_close_to_fp64_oracle(ora, nx, ny, kmult, bc):
    print(f"\n[case] f32 ext {nx}x{ny} kmult={kmult} bc={bc}, 200 steps vs FP64 oracle",
          flush=True)
    s = _setup(ora, nx, ny)
    dt = 0.9 * ora.dt_cfl_ext(s, kmult)
    k = ora.make_const(s, dt)
    k.kmult = kmult
    n = 200
    plan = ora.track_plan(1.0, dt, s.dom.Lx, s.dom.Ly, n)
    T0 = _init(nx, ny, 3 + nx)
    T32, st32 = run_f32_ext(k, plan, T0, n, bc=bc)
    T64, st64 = ora.run_plan(k, plan, T0, phys=1, bc=bc, T_bc=300.0, Tm=1609.0)
    err = float(np.abs(T32 - T64).max())
    pk = max(abs(a.tmax - b.tmax) for a, b in zip(st32, st64))
    print(f"  max|dT| {err:.3e} K, peak T differs by at most {pk:.3e} K, "
          f"peak {T64.max():.1f} K", flush=True)
    assert err <= 0.05
    assert pk <= 0.05
    for a, b in zip(st32, st64):
        assert abs(a.molten - b.molten) <= max(2, 0.01 * b.molten)
        for f in ("sum_src", "sum_loss", "sum_dT"):
            x, y = getattr(a, f), getattr(b, f)
            assert abs(x - y) <= 1e-3 * abs(y) + 1e-4 * nx * ny, (f, x, y)


@pytest.mark.parametrize("kmult", [1.0, 2.5])
def test_f32_ext_device_run_case0(ora, kmult):
    print(f"\n[case] DeviceRun FP32 ext case0 kmult={kmult}, whole run vs FP64", flush=True)
    mat, pr, dom = h.Material(), h.apply_case("0", h.Process()), h.Domain()
    dt = h.dt_cfl_ext(mat, dom, kmult)
    n = int(1.5e-3 / dt)
    k = h.make_const(mat, pr, dom, dt)
    k.kmult = kmult
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    out = {}
    for prec in (64, 32):
        with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan,
                         x_csv=x_csv, wrapped=wrapped, phys=1, precision=prec, batch=64) as r:
            assert r.advance(n) == n
            out[prec] = (r.download(), r.records(), r.summary())
    (T64, r64, s64), (T32, r32, s32) = out[64], out[32]
    e = np.abs(T32 - T64)
    err, p999, mean = float(e.max()), float(np.quantile(e, 0.999)), float(e.mean())
    at = np.unravel_index(int(e.argmax()), e.shape)
    dW = max(abs(a.W_um - b.W_um) for a, b in zip(r32, r64))
    dpk = max(abs(a.stats.tmax - b.stats.tmax) for a, b in zip(r32, r64))
    print(f"  {n} steps: max|dT| {err:.3e} K at (x={at[1]}, y={at[0]}, T={T64[at]:.1f} K), "
          f"99.9th percentile {p999:.3e} K, mean {mean:.3e} K; per-step peak T within "
          f"{dpk:.3e} K; peak T {s32[0]:.3f} / {s64[0]:.3f} K, peak W {s32[1]:.2f} / "
          f"{s64[1]:.2f} um, steady W_sub {s32[2]:.3f} / {s64[2]:.3f} um, max |dW| {dW:.2f} um",
          flush=True)
    assert all(x.has_col == y.has_col for x, y in zip(r32, r64))
    assert err <= 10.0 and p999 <= 1.0 and mean <= 0.01
    assert dpk <= 0.05
    assert abs(s32[0] - s64[0]) <= 0.05
    assert abs(s32[2] - s64[2]) <= 1.0
