"""The batched FP32 kernel on its own workload (256 instances, seed 0: sides
log-uniform in 256..2048, P, V, r0 drawn per instance, dx 10 um, dt 0.8x
the stability limit), autotuned per size class as the benchmark runs it:
a sample of instances (the smallest, the largest and some in between) is
bitwise the single-instance lpbf_launch_step_f32 field after 48 steps, and
the smallest ones stay within 0.05 K of the FP64 oracle."""
import numpy as np
import pytest
import torch

import bench
import hedgehog_lpbf_amd as h
from tests.test_gpu_f32 import run_f32

pytestmark = pytest.mark.gpu


def test_batch256_instances_equal_single_runs_and_oracle(ora):
    n = 48
    specs, dt = bench.batch_setup(n_steps=n, seed=0)
    assert len(specs) == 256
    cells = np.array([sp.k.nx * sp.k.ny for sp in specs])
    order = np.argsort(cells)
    sample = sorted(set(order[:4].tolist() + order[-2:].tolist() + order[::37].tolist()))
    print(f"\n[case] batch256 seed 0, {n} steps, sample {sample}", flush=True)
    b = h.F32Batch(specs, bc=0)
    ms = b.autotune(4)
    assert ms > 0
    b.enqueue(0, n)
    torch.cuda.synchronize()
    h.check()
    for i in sample:
        sp = specs[i]
        th1, _, _ = run_f32(sp.k, sp.plan, np.full((sp.k.ny, sp.k.nx), 300.0), n, 1, 0, 0,
                            T_ref=sp.T_ref)
        # the stored FP32 theta itself (T_ref + theta - T_ref in FP64 would
        # drop the far field's values near 1e-14 K)
        got = b.theta(i)[:, :sp.k.nx].cpu().numpy()
        assert np.array_equal(got, th1), (i, sp.k.nx, sp.k.ny)
    for i in order[:4]:
        sp = specs[i]
        k = h.StepConst.from_buffer_copy(bytes(sp.k))
        T_o, _ = ora.run_plan(k, sp.plan, np.full((sp.k.ny, sp.k.nx), 300.0))
        err = np.abs(b.field(i) - T_o).max()
        print(f"instance {i} ({sp.k.nx}x{sp.k.ny}): max|dT| vs FP64 oracle {err:.2e} K")
        assert err <= 0.05
    b.close()
