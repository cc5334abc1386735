"""Batched FP32 instances (include/lpbf_batch.h): instances of uneven size,
physics and tiling stepped by the multi-instance kernel must give, bit for
bit, the field each instance gets alone from lpbf_launch_step_f32 (whose
result does not depend on depth or band height); per-step statistics equal
the single-instance ones (integer parts exactly, FP64 energy sums to the
merge-order tolerance of the single-field invariance test); a captured
hipGraph replayed several times, with eager work between the replays,
equals the eager enqueue after every replay; the autotuner leaves the
fields as it found them."""
import numpy as np
import pytest
import torch

import hedgehog_lpbf_amd as h
from tests.test_gpu_f32 import _init, _setup, run_f32

pytestmark = pytest.mark.gpu

SHAPES = [  # nx, ny, size class, P, V, radiation
    (300, 200, 1, 285.0, 0.96, True),
    (513, 97, 1, 325.0, 1.2, True),
    (256, 256, 0, 200.0, 0.8, True),
    (1100, 150, 1, 400.0, 2.0, True),
    (250, 700, 0, 150.0, 0.5, True),
    (64, 40, 0, 285.0, 1.0, False),
]


def _specs(ora, n, T_ref=300.0):
    specs = []
    for q, (nx, ny, cls, P, V, rad) in enumerate(SHAPES):
        s = _setup(ora, nx, ny)
        s.proc.P, s.proc.V, s.proc.radiation = P, V, rad
        k = ora.make_const(s, 0.8 * ora.dt_cfl(s))
        plan = ora.track_plan(V, k.dt, s.dom.Lx, s.dom.Ly, n)
        specs.append(h.InstanceSpec(k=h.StepConst.from_buffer_copy(bytes(k)), plan=plan,
                                    T0=_init(nx, ny, 7 + q), T_ref=T_ref, size_class=cls))
    return specs


def _tol(nx, ny):
    return 1e-5 * nx * ny


@pytest.mark.parametrize("bc", [0, 1])
def test_batch_equals_single_instances(ora, bc):
    print(f"\n[case] f32 batch of {len(SHAPES)} instances bc={bc}, 37 steps", flush=True)
    n = 37
    specs = _specs(ora, n)
    b = h.F32Batch(specs, bc=bc)
    b.set_tiling(0, 2, 16)
    b.set_tiling(1, 4, 0)
    b.enqueue(0, 20)
    b.enqueue(20, n - 20)
    torch.cuda.synchronize()
    h.check()
    for i, sp in enumerate(specs):
        th1, st1, _ = run_f32(sp.k, sp.plan, sp.T0, n, 1, 0, bc, T_ref=sp.T_ref)
        got = b.field(i) - sp.T_ref
        assert np.array_equal(got.astype(np.float32), th1), i
        st = b.stats(i)
        for s, (x, y) in enumerate(zip(st, st1)):
            assert (x.tmax, x.arg, x.ymin, x.ymax, x.molten) == \
                (y.tmax, y.arg, y.ymin, y.ymax, y.molten), (i, s)
            for f in ("sum_dT", "sum_src", "sum_loss"):
                assert abs(getattr(x, f) - getattr(y, f)) <= _tol(sp.k.nx, sp.k.ny), (i, s, f)
    b.close()


def test_batch_close_to_fp64_oracle(ora):
    print("\n[case] f32 batch instance 0 vs FP64 oracle, 200 steps", flush=True)
    n = 200
    specs = _specs(ora, n)
    b = h.F32Batch(specs)
    b.enqueue(0, n)
    torch.cuda.synchronize()
    sp = specs[0]
    k = h.StepConst.from_buffer_copy(bytes(sp.k))
    T_o, _ = ora.run_plan(k, sp.plan, sp.T0)
    assert np.abs(b.field(0) - T_o).max() <= 0.05
    b.close()


def test_graph_replays_equal_eager(ora):
    print("\n[case] f32 batch hipGraph of 40 steps, 3 replays vs eager", flush=True)
    n = 40
    specs = _specs(ora, 3 * n)
    ga = h.F32Batch(specs)
    eb = h.F32Batch(specs)
    for b in (ga, eb):
        b.set_tiling(0, 3, 24)       # 40 steps at depth 3: 14 launches
        b.set_tiling(1, 4, 40)       # 10 launches
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ga.capture(0, n, s.cuda_stream)
    def same(tag):
        for i in range(len(specs)):
            assert torch.equal(ga.theta(i)[:, :specs[i].k.nx], eb.theta(i)[:, :specs[i].k.nx]), \
                (tag, i)

    for rep in range(3):
        with torch.cuda.stream(s):
            ga.launch_graph(s.cuda_stream)
        eb.enqueue(0, n)
        torch.cuda.synchronize()
        h.check()
        same(rep)
        for i in range(len(specs)):
            assert ga.launches(n) == eb.launches(n)
            sa, sb = ga.stats(i, 0, n), eb.stats(i, 0, n)
            assert [bytes(x) for x in sa] == [bytes(x) for x in sb], (rep, i)
        # eager work on the captured batch between replays: one single step
        # (an odd launch count) moves every field to its other buffer, so a
        # replay is refused until a second step moves them back
        with torch.cuda.stream(s):
            ga.enqueue(n + 2 * rep, 1, s.cuda_stream)
        eb.enqueue(n + 2 * rep, 1)
        with pytest.raises(h.LpbfError):
            with torch.cuda.stream(s):
                ga.launch_graph(s.cuda_stream)
        with torch.cuda.stream(s):
            ga.enqueue(n + 2 * rep + 1, 1, s.cuda_stream)
        eb.enqueue(n + 2 * rep + 1, 1)
        torch.cuda.synchronize()
        h.check()
        same((rep, "eager"))
    ga.close()
    eb.close()


def test_autotune_keeps_fields_and_picks_valid_tiling(ora):
    print("\n[case] f32 batch autotune", flush=True)
    specs = _specs(ora, 16)
    b = h.F32Batch(specs)
    b.enqueue(0, 1)                          # one launch: every field now in theta1
    torch.cuda.synchronize()
    assert all(h.lib().lpbf_f32_batch_current(b._h, i) == 1 for i in range(len(specs)))
    before = [b.field(i) for i in range(len(specs))]
    ms = b.autotune(4)
    assert ms > 0
    for c in (0, 1):
        d, r, m = b.tiling(c)
        assert 2 <= d <= 4 and (r == 0 or r >= 2 * d) and m == 3
    for i in range(len(specs)):
        assert np.array_equal(b.field(i), before[i]), i
    b.enqueue(1, 5)
    torch.cuda.synchronize()
    h.check()
    b.close()
