"""Pins the CPU reference of the surrogate network (oracle/surrogate_ref.py):
float64 gradients against central finite differences on a tiny window, the
conv against a direct per-pixel loop, the gelu derivative against finite
differences, and the bf16 rounding against known bit patterns."""
import numpy as np

from oracle import surrogate_ref as S


def test_parameter_count():
    assert S.N_PARAMS == 38465
    th = S.init_params(3)
    assert th.shape == (38465,)
    assert np.allclose(S.pack(S.unpack(th)), th)


def test_conv_equals_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 3))
    W = rng.normal(size=(9, 3, 4))
    b = rng.normal(size=4)
    z = S.conv(x, W, b)
    ref = np.zeros((2, 5, 6, 4))
    for bb in range(2):
        for y in range(5):
            for xx in range(6):
                acc = b.copy()
                for t in range(9):
                    yy, xs = y + t // 3 - 1, xx + t % 3 - 1
                    if 0 <= yy < 5 and 0 <= xs < 6:
                        acc += x[bb, yy, xs] @ W[t]
                ref[bb, y, xx] = acc
    assert np.allclose(z, ref, atol=1e-12)


def test_gelu_grad_finite_difference():
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (S.gelu(z + h) - S.gelu(z - h)) / (2 * h)
    assert np.allclose(S.gelu_grad(z), fd, atol=1e-8)
    assert abs(S.gelu(0.0)) == 0.0 and abs(S.gelu(10.0) - 10.0) < 1e-12


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265])
    r = S.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even
    assert r[2] == 1.0 + 4 * 2 ** -8               # tie to even (up)
    assert r[3] == -2.5 and abs(r[4] - 3.140625) == 0.0


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    th = S.init_params(5)
    th[-289:] = rng.normal(0, 0.05, 289)          # a non-trivial output layer
    x = rng.normal(size=(1, 6, 5, 4)) * 0.5
    tgt = rng.normal(size=(1, 6, 5)) * 0.1
    loss, g, _ = S.loss_and_grad(th, x, tgt)
    idx = np.concatenate([rng.choice(1152, 6, replace=False),               # W1
                          1152 + np.arange(3),                              # b1
                          1184 + rng.choice(9216, 6, replace=False),        # W2
                          1184 + 3 * 9248 + rng.choice(9216, 4, replace=False),   # W5
                          38465 - 289 + rng.choice(288, 4, replace=False),  # W6
                          [38464]])                                         # b6
    h = 1e-5
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (S.loss_and_grad(tp, x, tgt)[0] - S.loss_and_grad(tm, x, tgt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
