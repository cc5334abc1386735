"""lpbf_launch_step_f32_noalloc (statistics records in the library's fixed
device scratch) against lpbf_launch_step_f32 (caller workspace): the same
kernels, so fields and per-step statistics must agree bit for bit.  Two
calls in a row share the scratch, the second on a smaller plate (other
depth, band height and edge rule), so stale records of the first cannot
leak into the second's statistics."""
import numpy as np
import pytest
import torch

import hedgehog_lpbf_amd as h
from tests.devutil import dev, dev_bytes, kconst, plan_dev, stats_np, stream, sync_check
from tests.test_gpu_f32 import _init, _setup

pytestmark = pytest.mark.gpu


def _pair(ora, nx, ny, depth, rows, bc, steps):
    """One launch sequence of `steps` steps through each entry point."""
    lib = h.lib()
    s = _setup(ora, nx, ny)
    k = ora.make_const(s, ora.dt_cfl(s))
    plan = ora.track_plan(3.0, k.dt, s.dom.Lx, s.dom.Ly, steps)
    kk = kconst(k)
    ld = (nx + 3) // 4 * 4
    th0 = np.zeros((ny, ld), dtype=np.float32)
    th0[:, :nx] = (_init(nx, ny, nx + ny) - 300.0).astype(np.float32)
    pdev = plan_dev(plan)
    out = []
    for noalloc in (False, True):
        a = dev(th0)
        b = torch.zeros_like(a)
        st = dev_bytes(56 * steps)
        work = None if noalloc else dev_bytes(lib.lpbf_f32_workspace_bytes(nx, ny, depth, rows))
        s0 = 0
        while s0 < steps:
            d = min(depth, steps - s0)
            if noalloc:
                lib.lpbf_launch_step_f32_noalloc(a.data_ptr(), b.data_ptr(), ld, kk, 300.0, bc,
                                                 300.0, 1609.0, pdev.data_ptr(), s0, d, rows,
                                                 st.data_ptr() + 56 * s0, stream())
            else:
                lib.lpbf_launch_step_f32(a.data_ptr(), b.data_ptr(), ld, kk, 300.0, bc, 300.0,
                                         1609.0, pdev.data_ptr(), s0, d, rows,
                                         st.data_ptr() + 56 * s0, work.data_ptr(), stream())
            a, b = b, a
            s0 += d
        sync_check()
        out.append((a[:, :nx].cpu().numpy(), [bytes(x) for x in stats_np(st, steps)]))
        del a, b, st, work
    return out


@pytest.mark.parametrize("first,second", [
    ((8192, 8192, 4, 46, 0, 8), (1003, 700, 3, 7, 1, 9)),
    ((2000, 1500, 4, 0, 1, 6), (37, 5, 2, 0, 0, 5)),
])
def test_noalloc_equals_workspace_entry_point(ora, first, second):
    for nx, ny, depth, rows, bc, steps in (first, second):
        print(f"\n[case] f32 noalloc vs workspace {nx}x{ny} depth {depth} rows {rows} bc {bc} "
              f"{steps} steps", flush=True)
        need = h.lib().lpbf_f32_workspace_bytes(nx, ny, depth, rows)
        assert 0 < need <= h._lib.LPBF_F32_NOALLOC_SCRATCH_BYTES
        (fa, sa), (fb, sb) = _pair(ora, nx, ny, depth, rows, bc, steps)
        assert np.array_equal(fa, fb)
        assert sa == sb
