"""Surrogate kernels (include/lpbf_surrogate.h) against the CPU reference of
the network (oracle/surrogate_ref.py, which rounds to bf16 where the GPU
stores bf16): forward output, loss and every layer's gradient; Adam against
its formula; window gather / target / blend against numpy; and what the
trained network gains over the method without it (persistence,
T_{t+16} = T_t) on held-out windows of FD runs.

Tolerances: the GPU accumulates in FP32 in a different order than the FP64
reference, which can move a stored bf16 value by one unit (2^-8 relative)
now and then; forward outputs agree to 1e-2 of their range, the loss to
1e-3 relative and each layer's gradient to 3e-2 in relative L2 norm."""
import numpy as np
import pytest
import torch

import hedgehog_lpbf_amd as h
from hedgehog_lpbf_amd import surrogate as sg
from oracle import surrogate_ref as S

pytestmark = pytest.mark.gpu


def _inputs(B, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:128, 0:128]
    xs = np.zeros((B, 128, 128, 4))
    for b in range(B):
        cx, cy = 50 + 10 * b, 64
        hot = 2.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 300.0)
        xs[b, ..., 0] = hot + 0.01 * rng.normal(size=(128, 128))
        xs[b, ..., 1] = np.exp(-((x - cx - 3) ** 2 + (y - cy) ** 2) / 60.0)
        xs[b, ..., 2] = 0.7
        xs[b, ..., 3] = 0.55
    tgt = 0.05 * np.tanh(xs[..., 0]) - 0.02 * xs[..., 1] + 0.001 * rng.normal(size=(B, 128, 128))
    return S.bf16(xs), tgt.astype(np.float32).astype(np.float64)


def _x0(xs):
    B = xs.shape[0]
    x0 = torch.zeros((B, 128, 128, 32), dtype=torch.bfloat16)
    x0[..., :4] = torch.from_numpy(xs.astype(np.float32)).to(torch.bfloat16)
    return x0.cuda()


def _params(seed):
    th = S.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    th[-289:] = rng.normal(0, 0.05, 289)
    for q in range(5):                       # small non-zero biases
        o = 1184 + (q - 1) * 9248 + 9216 if q else 1152
        th[o:o + 32] = rng.normal(0, 0.05, 32)
    return th.astype(np.float32).astype(np.float64)


def test_forward_matches_reference():
    print("\n[case] surrogate forward, 2 windows", flush=True)
    th = _params(1)
    xs, _ = _inputs(2, 2)
    m = sg.Surrogate(params=th.astype(np.float32))
    y = m.forward(_x0(xs))
    torch.cuda.synchronize()
    h.check()
    ref = S.forward(th, xs, bf=True)
    got = y.cpu().numpy().astype(np.float64)
    scale = np.abs(ref).max()
    assert scale > 1e-3
    assert np.abs(got - ref).max() <= 1e-2 * scale, (np.abs(got - ref).max(), scale)
    assert np.abs(got - ref).mean() <= 1e-3 * scale


def test_loss_and_gradients_match_reference():
    print("\n[case] surrogate loss + gradients, 1 window", flush=True)
    th = _params(3)
    xs, tgt = _inputs(1, 4)
    m = sg.Surrogate(params=th.astype(np.float32))
    t = torch.from_numpy(tgt.astype(np.float32)).cuda()
    loss, grad = m.loss_grad(_x0(xs), t)
    torch.cuda.synchronize()
    h.check()
    l_ref, g_ref, _ = S.loss_and_grad(th, xs, tgt, bf=True)
    assert abs(float(loss.item()) - l_ref) <= 1e-3 * l_ref
    g = grad.cpu().numpy().astype(np.float64)
    bounds = [0, 1152, 1184]
    for q in range(4):
        bounds += [bounds[-1] + 9216, bounds[-1] + 9248]
    bounds += [bounds[-1] + 288, bounds[-1] + 289]
    assert bounds[-1] == 38465
    for a, b in zip(bounds[:-1], bounds[1:]):
        err = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        assert err <= 3e-2, (a, b, err)


def test_adam_matches_formula():
    print("\n[case] surrogate adam", flush=True)
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=38465).astype(np.float32)
    m = sg.Surrogate(params=p0)
    p = p0.astype(np.float64)
    mm = np.zeros_like(p)
    vv = np.zeros_like(p)
    for t in range(1, 4):
        g = rng.normal(size=38465).astype(np.float32)
        m.grad.copy_(torch.from_numpy(g))
        m.adam(lr=1e-2)
        mm = 0.9 * mm + 0.1 * g
        vv = 0.999 * vv + 0.001 * g.astype(np.float64) ** 2
        p -= 1e-2 * (mm / (1 - 0.9 ** t)) / (np.sqrt(vv / (1 - 0.999 ** t)) + 1e-8)
    torch.cuda.synchronize()
    assert np.abs(m.params.cpu().numpy() - p).max() <= 1e-5


def test_gather_target_apply():
    print("\n[case] surrogate windows gather/target/apply", flush=True)
    ny, nx, ld = 200, 300, 300
    rng = np.random.default_rng(6)
    th0 = (rng.uniform(0, 2000, (ny, ld))).astype(np.float32)
    th1 = th0 + rng.normal(0, 5, (ny, ld)).astype(np.float32)
    d0, d1 = torch.from_numpy(th0).cuda(), torch.from_numpy(th1).cuda()
    plan = np.zeros((40, 3))
    plan[:, 0] = 1e-5 * (100 + np.arange(40) * 0.7)
    plan[:, 1] = 1e-5 * 90
    plan[:, 2] = 1.0
    plan[20:, 2] = 0.0
    pd = torch.from_numpy(plan).cuda()
    wins = [sg.SgWindow(d0.data_ptr(), d1.data_ptr(), pd.data_ptr(), 300.0, 1e-5, 1e-5,
                        1 / (50e-6) ** 2, 250.0, 1.2, ld, nx, ny, 37, 50, 0),
            sg.SgWindow(d0.data_ptr(), d1.data_ptr(), pd.data_ptr(), 300.0, 1e-5, 1e-5,
                        1 / (50e-6) ** 2, 300.0, 0.9, ld, nx, ny, 172, 72, 10)]
    wd = sg.windows_buffer(wins)
    x0 = torch.empty((2, 128, 128, 32), dtype=torch.bfloat16, device="cuda")
    T0 = torch.empty((2, 128, 128), dtype=torch.float32, device="cuda")
    sg.gather(wd, 2, x0, T0)
    x = x0.float().cpu().numpy()
    yy, xx = np.mgrid[0:128, 0:128]
    for b, w in enumerate(wins):
        sl = th0[w.oy:w.oy + 128, w.ox:w.ox + 128]
        assert np.array_equal(T0[b].cpu().numpy(), sl)
        assert np.array_equal(x[b, ..., 0], S.bf16(sl / np.float32(3000.0)))
        src = np.zeros((128, 128))
        for s in range(16):
            xl, yl, q = plan[w.step + s]
            if q:
                r2 = ((w.ox + xx) * 1e-5 - xl) ** 2 + ((w.oy + yy) * 1e-5 - yl) ** 2
                src += q * np.exp(-2 * r2 * w.inv_r0_sq)
        assert np.abs(x[b, ..., 1] - S.bf16(src / 16)).max() <= 2 ** -8 * max(src.max() / 16, 1e-30)
        assert np.all(x[b, ..., 2] == S.bf16(np.float32(w.P / 400)))
        assert np.all(x[b, ..., 3] == S.bf16(np.float32(w.V / 2)))
        assert not x[b, ..., 4:].any()
    # target from the field 16 steps later (here: th1)
    wins_t = [sg.SgWindow.from_buffer_copy(bytes(w)) for w in wins]
    for w in wins_t:
        w.theta = d1.data_ptr()
    wdt = sg.windows_buffer(wins_t)
    tg = torch.empty_like(T0)
    sg.target(wdt, 2, T0, tg)
    for b, w in enumerate(wins):
        ref = (th1[w.oy:w.oy + 128, w.ox:w.ox + 128] - th0[w.oy:w.oy + 128, w.ox:w.ox + 128]) / \
            np.float32(3000.0)
        assert np.allclose(tg[b].cpu().numpy(), ref, rtol=0, atol=1e-7)
    # blend: prediction = T0 + 3000 y written into d1 with an 8-cell ramp
    y = torch.from_numpy(rng.normal(0, 1e-3, (2, 128, 128)).astype(np.float32)).cuda()
    md = torch.zeros(2, dtype=torch.float32, device="cuda")
    skip = torch.tensor([0, 1], dtype=torch.int32, device="cuda")
    before = d1.cpu().numpy().copy()
    sg.apply(wd, 2, T0, y, 8, skip, md)
    torch.cuda.synchronize()
    h.check()
    after = d1.cpu().numpy()
    w = wins[0]
    pred = T0[0].cpu().numpy() + np.float32(3000.0) * y[0].cpu().numpy()
    cur = before[w.oy:w.oy + 128, w.ox:w.ox + 128]
    e = np.minimum(np.minimum(yy, xx), np.minimum(127 - yy, 127 - xx))
    a = np.minimum(1.0, e / 8.0).astype(np.float32)
    assert np.allclose(after[w.oy:w.oy + 128, w.ox:w.ox + 128], a * pred + (1 - a) * cur,
                       rtol=0, atol=1e-3)
    mask = np.ones_like(before, dtype=bool)
    mask[w.oy:w.oy + 128, w.ox:w.ox + 128] = False
    assert np.array_equal(after[mask], before[mask])           # window 1 skipped
    for b in range(2):
        wb = wins[b]
        p = T0[b].cpu().numpy() + np.float32(3000.0) * y[b].cpu().numpy()
        ref = np.abs(p - before[wb.oy:wb.oy + 128, wb.ox:wb.ox + 128]).max()
        assert abs(float(md[b]) - ref) <= 1e-3 * ref


def test_trained_network_beats_persistence():
    print("\n[case] surrogate: FD data (8 runs x 1024 steps), 300 Adam steps", flush=True)
    spec = sg.DataSpec(n_runs=8, nx=384, ny=256, steps=1024, n_tracks=2, seed=11)
    xs, T0s, tg = sg.generate(spec, log=print)
    N = xs.shape[0]
    assert N == 8 * (1024 // 16 - 1)
    rng = np.random.default_rng(0)
    perm = rng.permutation(N)
    test_idx, train_idx = perm[:64], perm[64:]
    m = sg.Surrogate(seed=0)
    before = sg.evaluate(m, xs, T0s, tg, test_idx, batch=32)
    tr = torch.as_tensor(train_idx, device="cuda")
    losses = sg.train(m, xs[tr], tg[tr], iters=300, batch=32, lr=1e-3, log=print, log_every=100)
    after = sg.evaluate(m, xs, T0s, tg, test_idx, batch=32)
    print(f"[surrogate] held-out rms K: persistence {after['rms_persistence']:.2f}, "
          f"untrained {before['rms_net']:.2f}, trained {after['rms_net']:.2f}; "
          f"max K: persistence {after['max_persistence']:.1f}, trained {after['max_net']:.1f}",
          flush=True)
    assert losses[-1][1] < losses[0][1]
    assert after["rms_net"] < 0.5 * after["rms_persistence"]
