"""Surrogate kernels (include/lpbf_surrogate.h) against the CPU reference of
the network (oracle/surrogate_ref.py, which rounds to bf16 where the GPU
stores bf16): forward output, loss and every layer's gradient; Adam against
its formula; and what the trained network gains over the method without it
(the two-level step's prolonged coarse increment alone) on held-out windows
of FD runs.

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


def test_forward_and_gradients_are_deterministic():
    """Repeated passes over the same inputs give the same bits: every row a
    conv workgroup reads from its LDS ring has landed before it is read
    (including the rows past a window's last, which have no transfer)."""
    print("\n[case] surrogate determinism, 8 windows x 4 passes", flush=True)
    th = _params(3)
    xs, tgt = _inputs(8, 4)
    m = sg.Surrogate(params=th.astype(np.float32))
    x0 = _x0(xs)
    ys = [m.forward(x0).cpu().numpy() for _ in range(4)]
    h.check()
    for y in ys[1:]:
        assert np.array_equal(y, ys[0])
    t = torch.from_numpy(tgt.astype(np.float32)).cuda()
    gs = []
    for _ in range(3):
        _, grad = m.loss_grad(x0, t)
        gs.append(grad.cpu().numpy().copy())
    h.check()
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])


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


def test_trained_network_beats_the_two_level_step_without_it():
    """Training on the two-level run's pairs (lpbf_mr_gather / lpbf_mr_target
    over FD runs and their coarse companions): the held-out residual error of
    the trained network is well below that of the method without it (the
    prolonged coarse increment alone, residual 0)."""
    print("\n[case] surrogate: FD data (8 runs x 1024 steps), 300 Adam steps", flush=True)
    spec = sg.DataSpec(n_runs=8, nx=384, ny=256, steps=1024, n_tracks=2, seed=11)
    xs, Ds, tg = sg.generate(spec, log=print)
    N = xs.shape[0]
    assert N == 8 * (1024 // 16 - 1)
    rng = np.random.default_rng(0)
    perm = rng.permutation(N)
    test_idx, train_idx = perm[:64], perm[64:]
    m = sg.Surrogate(seed=0)
    before = sg.evaluate(m, xs, tg, test_idx, batch=32)
    tr = torch.as_tensor(train_idx, device="cuda")
    losses = sg.train(m, xs[tr], tg[tr], iters=300, batch=32, lr=1e-3, log=print, log_every=100)
    after = sg.evaluate(m, xs, tg, test_idx, batch=32)
    print(f"[surrogate] held-out rms K: without network {after['rms_without_net']:.3f}, "
          f"untrained {before['rms_net']:.3f}, trained {after['rms_net']:.3f}; max K: without "
          f"{after['max_without_net']:.1f}, trained {after['max_net']:.1f}", flush=True)
    assert losses[-1][1] < losses[0][1]
    assert after["rms_without_net"] > 0.1
    assert after["rms_net"] < 0.5 * after["rms_without_net"]
