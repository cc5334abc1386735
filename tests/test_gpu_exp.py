"""The one libm function at the FP64 parity boundary: the device exp the step
kernels call (lpbf_probe_exp_f64) against the host's glibc exp (Python's
math.exp), in units in the last place, over the arguments the step produces
(the Gaussian exponent -2 r^2 / r0^2 on the reference case0 grid for several
beam positions, every value at or above the -746 cut-off) and uniform
samples of [-746, 0].  The reference's GPU port has the same boundary
(src/gpu/kernels.cu:115 vs include/step_task.hpp:113) and no test of it."""
import math

import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from tests.devutil import dev, stream, sync_check

pytestmark = pytest.mark.gpu


def _ulps(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return np.abs(ia - ib)


def _args():
    mat, pr, dom = h.Material(), h.apply_case("0", h.Process()), h.Domain()
    k = h.make_const(mat, pr, dom, h.dt_cfl(mat, dom))
    y, x = np.mgrid[0:dom.ny, 0:dom.nx]
    xs, ys = x * k.dx, y * k.dy
    out = []
    for fx in (0.2, 0.37, 0.5, 0.81):
        xl, yl = fx * dom.Lx, 0.5 * dom.Ly
        r2 = (xs - xl) * (xs - xl) + (ys - yl) * (ys - yl)
        a = -2.0 * r2 * k.inv_r0_sq
        out.append(a[a >= -746.0])
    rng = np.random.default_rng(7)
    out.append(-746.0 * rng.random(200_000))
    out.append(-1e-3 * rng.random(20_000))
    return np.concatenate(out)


def test_device_exp_within_one_ulp_of_glibc():
    print("\n[case] device exp vs glibc exp", flush=True)
    x = _args()
    xd, y = dev(x), dev(np.zeros_like(x))
    h.lib().lpbf_probe_exp_f64(xd.data_ptr(), y.data_ptr(), x.size, stream())
    sync_check()
    got = y.cpu().numpy()
    ref = np.array([math.exp(v) for v in x.tolist()])
    u = _ulps(got, ref)
    print(f"[exp] {x.size} arguments: {np.mean(u == 0) * 100:.2f}% identical, "
          f"{np.mean(u == 1) * 100:.2f}% one ulp apart, max {u.max()} ulp", flush=True)
    assert np.all(np.isfinite(got)) and np.all(got >= 0.0)
    assert u.max() <= 1
