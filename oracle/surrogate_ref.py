"""CPU reference of the melt-pool surrogate network (test infrastructure only).

Only tests/ and __graft_entry__.smoke() use this module, to check the GPU
kernels of hedgehog_lpbf_amd/csrc/lpbf_surrogate.hip.  The reference project
has no surrogate; this restates the network the library implements:

  input  x[b, y, x, 4]: (T_t - Tamb)/3000, laser source map, P/400, V/2
  L1..L5 z = conv3x3(a_prev) + bias (zero padding), a = gelu(z), 32 channels
  L6     y = conv3x3(a5) + bias, one channel: the predicted (T_{t+16}-T_t)/3000
  loss   mean over pixels of (y - target)^2

conv3x3: z[p, co] = b[co] + sum_t sum_ci x[p + off(t), ci] * W[t, ci, co] with
tap t = (dy+1)*3 + (dx+1), off(t) = (dy, dx), and x = 0 outside the window.
gelu is the tanh form 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi).

Parameter vector (38,465 values): W1[9,4,32], b1[32], then for L2..L5
W[9,32,32], b[32], then W6[9,32], b6[1].

With bf16=True the forward/backward round where the GPU stores bf16 (the
input, weights, z and a of L1..L5, and the backward dz of L1..L5), so the
GPU result differs from this one only by accumulation order.  Pinned by
finite differences in float64 (tests/test_oracle_surrogate.py).
"""
from __future__ import annotations

import numpy as np

C_IN, C = 4, 32
LAYERS = [(C_IN, C), (C, C), (C, C), (C, C), (C, C), (C, 1)]
N_PARAMS = sum(9 * ci * co + co for ci, co in LAYERS)
GELU_C = float(np.sqrt(2.0 / np.pi))


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Flat parameter vector -> [(W[9, ci, co], b[co]), ...]."""
    out, o = [], 0
    for ci, co in LAYERS:
        W = theta[o:o + 9 * ci * co].reshape(9, ci, co)
        o += 9 * ci * co
        b = theta[o:o + co]
        o += co
        out.append((W, b))
    assert o == N_PARAMS
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(seed=0):
    """He-style initialisation; the output layer starts at zero increment."""
    rng = np.random.default_rng(seed)
    layers = []
    for q, (ci, co) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (9 * ci))
        W = rng.normal(0.0, std, (9, ci, co))
        if q == len(LAYERS) - 1:
            W *= 0.1
        layers.append((W, np.zeros(co)))
    return pack(layers)


def _shift(x, dy, dx):
    """xs[b, y, x] = x[b, y+dy, x+dx] (0 outside)."""
    B, H, W_, Cc = x.shape
    out = np.zeros_like(x)
    ys0, ys1 = max(0, -dy), min(H, H - dy)
    xs0, xs1 = max(0, -dx), min(W_, W_ - dx)
    out[:, ys0:ys1, xs0:xs1] = x[:, ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
    return out


def conv(x, W, b):
    z = np.zeros(x.shape[:3] + (W.shape[2],))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        z += np.einsum("bhwc,cd->bhwd", _shift(x, dy, dx), W[t])
    return z + b


def conv_bwd(x, W, dz):
    """Gradients of conv: (dx, dW, db)."""
    dW = np.zeros_like(W)
    dx_ = np.zeros_like(x)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        dW[t] = np.einsum("bhwc,bhwd->cd", _shift(x, dy, dx), dz)
        # x[p + off] feeds z[p]: the adjoint sends dz[p] back to p + off
        dx_ += _shift(np.einsum("bhwd,cd->bhwc", dz, W[t]), -dy, -dx)
    return dx_, dW, dz.sum(axis=(0, 1, 2))


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = GELU_C * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3 * 0.044715 * z * z)


def forward(theta, x, bf=False, keep=False):
    layers = unpack(theta)
    r = bf16 if bf else (lambda v: v)
    a = r(x)
    cache = []
    for q, (W, b) in enumerate(layers):
        Wq = r(W) if q < 5 else r(W)
        z = conv(a, Wq, b)
        if q == 5:
            cache.append((a, Wq, None))
            return (z[..., 0], cache) if keep else z[..., 0]
        zs = r(z)                        # stored pre-activation
        a_next = r(gelu(z))              # activation from the unrounded z
        cache.append((a, Wq, zs))
        a = a_next


def loss_and_grad(theta, x, target, bf=False):
    """(loss, grad[N_PARAMS], y)."""
    y, cache = forward(theta, x, bf=bf, keep=True)
    r = bf16 if bf else (lambda v: v)
    n = y.size
    d = y - target
    loss = float(np.mean(d * d))
    dz = (2.0 / n) * d[..., None]
    grads = [None] * 6
    for q in range(5, -1, -1):
        a, Wq, _ = cache[q]
        da, dW, db = conv_bwd(a, Wq, dz)
        grads[q] = (dW, db)
        if q > 0:
            zs_prev = cache[q - 1][2]
            dz = r(da * gelu_grad(zs_prev))
    return loss, pack(grads), y
