"""FP64 steps with temporal blocking (lpbf_launch_step_f64_blocked) against
`depth` single steps of lpbf_launch_step_ex on the same device: every cell of
every step is the same expression on the same operands, so the fields must
be bit-identical (exp included, both run the device exp), for both physics,
both boundary conditions, depths 1-4, band heights that split the plate
unevenly, and ragged plate sizes.  Per-step statistics: max, index, molten
rows and count exact; the FP64 energy sums to 1e-12 relative (they are added
in another order).  Against the CPU oracle the bar is the usual 1e-9 K."""
import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from tests.devutil import dev, dev_bytes, kconst, plan_dev, stats_np, stream, sync_check

SIZES = [(37, 5), (131, 67), (300, 133), (2, 9), (129, 2), (240, 120), (1100, 150)]


def _setup(ora, nx, ny, phys, seed):
    s = ora.defaults("0")
    s.dom.nx, s.dom.ny = nx, ny
    k = ora.make_const(s, ora.dt_cfl(s) if phys == 0 else ora.dt_cfl_ext(s, 1.8))
    k.kmult = 1.8 if phys else 1.0
    rng = np.random.default_rng(seed)
    T = 300.0 + 2300.0 * rng.random((ny, nx))
    n = 8
    plan = np.zeros((n, 3))
    plan[:, 0] = (0.3 + 0.05 * np.arange(n)) * s.dom.Lx
    plan[:, 1] = 0.45 * s.dom.Ly
    plan[:, 2] = 1.0
    plan[5, 2] = 0.0            # a step with the beam off
    return k, T, plan


def _single_steps(k, T, plan, phys, bc, s0, n):
    nx, ny = k.nx, k.ny
    nb = h.lib().lpbf_step_ex_nblocks(nx, ny)
    blk = dev_bytes(nb * 56)
    outs = dev_bytes(56 * n)
    a, b = dev(T), dev(np.zeros_like(T))
    st = stream()
    kk = kconst(k)
    fields = []
    for s in range(s0, s0 + n):
        kk.x_l, kk.y_l, kk.q_scale = plan[s]
        h.lib().lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kk, phys, bc, 300.0, 1609.0,
                                    blk.data_ptr(), st)
        h.lib().lpbf_launch_stats_finalize(blk.data_ptr(), nb, nx, ny,
                                           outs.data_ptr() + 56 * (s - s0), st)
        a, b = b, a
        fields.append(a)
    sync_check()
    return [f.cpu().numpy() for f in fields], stats_np(outs, n)


def _blocked(k, T, pd, phys, bc, s0, depth, rows):
    nx, ny = k.nx, k.ny
    work = dev_bytes(h.lib().lpbf_f64_blocked_workspace_bytes(nx, ny, depth, rows))
    stats = dev_bytes(56 * depth)
    a, b = dev(T), dev(np.zeros_like(T))
    h.lib().lpbf_launch_step_f64_blocked(a.data_ptr(), b.data_ptr(), kconst(k), phys, bc, 300.0,
                                         1609.0, pd.data_ptr(), s0, depth, rows,
                                         stats.data_ptr(), work.data_ptr(), stream())
    sync_check()
    return b.cpu().numpy(), stats_np(stats, depth)


def _same_stats(got, ref):
    for g, r in zip(got, ref):
        assert (g.tmax, g.arg, g.ymin, g.ymax, g.molten) == (r.tmax, r.arg, r.ymin, r.ymax,
                                                             r.molten)
        for f in ("sum_dT", "sum_src", "sum_loss"):
            assert getattr(g, f) == pytest.approx(getattr(r, f), rel=1e-12, abs=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("phys,bc", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_blocked_equals_single_steps(ora, nx, ny, phys, bc):
    print(f"\n[case] f64 blocked nx={nx} ny={ny} phys={phys} bc={bc}", flush=True)
    k, T, plan = _setup(ora, nx, ny, phys, nx * 31 + ny)
    pd = plan_dev(plan)
    for depth, rows, s0 in ((1, 0, 0), (2, 3, 1), (3, 0, 2), (4, 5, 0), (4, 0, 3)):
        ref_fields, ref_stats = _single_steps(k, T, plan, phys, bc, s0, depth)
        got, st = _blocked(k, T, pd, phys, bc, s0, depth, rows)
        same = got.view(np.uint64) == ref_fields[-1].view(np.uint64)
        assert same.all(), (depth, rows, int((~same).sum()), np.argwhere(~same)[:5].tolist())
        _same_stats(st, ref_stats)


@pytest.mark.gpu
@pytest.mark.parametrize("phys", [0, 1])
def test_blocked_against_oracle(ora, phys):
    print(f"\n[case] f64 blocked vs oracle phys={phys}", flush=True)
    k, T, plan = _setup(ora, 300, 133, phys, 99)
    got, st = _blocked(k, T, plan_dev(plan), phys, 0, 0, 4, 0)
    ref = T.copy()
    for s in range(4):
        kk = ora.StepConst.from_buffer_copy(bytes(k))
        kk.x_l, kk.y_l, kk.q_scale = plan[s]
        ref = ora.step(ref, kk, phys)
    assert np.abs(got - ref).max() <= 1e-9


def test_blocked_rejects_bad_arguments_without_launching():
    lib = h.lib()
    lib.lpbf_get_last_error()
    k = h.StepConst()
    k.nx, k.ny = 64, 64
    fake = 1 << 20     # never dereferenced: every call below fails validation
    for depth, rows, tin, tout in ((0, 0, fake, fake * 2), (5, 0, fake, fake * 2),
                                   (2, -1, fake, fake * 2), (2, 0, fake, fake)):
        lib.lpbf_launch_step_f64_blocked(tin, tout, k, 0, 0, 300.0, 1609.0, fake, 0, depth, rows,
                                         None, None, None)
        assert lib.lpbf_get_last_error() == h._lib.LPBF_ERR_INVALID_ARGUMENT
    assert lib.lpbf_f64_blocked_workspace_bytes(64, 64, 5, 0) == 0
    assert lib.lpbf_f64_blocked_workspace_bytes(64, 64, 2, 0) > 0
