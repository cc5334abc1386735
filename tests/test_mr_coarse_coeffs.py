"""The two-level run's coarse-step coefficients (lpbf_mr_coarse_coeffs,
host arithmetic only).

Pins:
  * without radiation the loss is linear and four fine steps have a closed
    form: Phi^4(theta) = (1 - a1)^4 theta - a0 (1 + q + q^2 + q^3), q = 1 - a1;
  * with radiation the quartic equals Phi iterated four times (brute force,
    FP64) to 1e-5 K over 0..3000 K of excess (the terms above theta^4 that
    the quartic drops);
  * the diffusion numbers are the fine ones times (1 - a1)^3;
  * on a hot ridge (the cooling trail of a track), a coarse field stepped with
    these coefficients stays within 0.2 K of the 2x2 mean of the fine FD
    field, where one plain step of four times the length drifts by over 1 K.
"""
import numpy as np
import pytest

import hedgehog_lpbf_amd as h
from hedgehog_lpbf_amd import surrogate as sg

SIGMA = 5.670374419e-8


def _fine_const(radiation=True):
    mat, pr, dom, k, plan = sg.run_setup(sg.RunSpec(nx=256, ny=256, steps=16, n_tracks=1))
    k.radiation = bool(radiation)
    return k


def _loss(k, T_ref, th):
    """dt * loss(theta) in K, from the plate physics (convection, substrate,
    radiation) in FP64."""
    T = T_ref + th
    q = k.hConv * (T - k.Tamb) + (k.K / k.dSub) * (T - k.Tsub)
    if k.radiation:
        q = q + k.eps * SIGMA * (T ** 4 - k.Tamb ** 4)
    return k.dt * q * k.inv_h / k.rho_cp


def _coeffs(k, T_ref):
    out = np.zeros(7)
    assert h.lib().lpbf_mr_coarse_coeffs(k, T_ref, out.ctypes.data) == 0
    return out


def _poly(p, th):
    return p[0] + th * (p[1] + th * (p[2] + th * (p[3] + th * p[4])))


def test_linear_loss_closed_form():
    k = _fine_const(radiation=False)
    T_ref = 350.0                                    # a0 != 0
    c = _coeffs(k, T_ref)
    a0 = _loss(k, T_ref, 0.0)
    a1 = _loss(k, T_ref, 1.0) - a0
    q = 1.0 - a1
    cx = k.alpha * k.dt * k.inv_dx2 * q ** 3
    cy = k.alpha * k.dt * k.inv_dy2 * q ** 3
    assert c[0] == pytest.approx(cx, rel=1e-12)
    assert c[1] == pytest.approx(cy, rel=1e-12)
    assert c[2] == pytest.approx(-a0 * (1 + q + q * q + q ** 3), rel=1e-10)
    assert c[3] == pytest.approx(q ** 4 - 2 * (cx + cy), rel=1e-12)
    assert np.all(np.abs(c[4:]) < 1e-30)


def test_radiation_quartic_equals_four_fine_loss_maps():
    k = _fine_const(radiation=True)
    T_ref = 300.0
    c = _coeffs(k, T_ref)
    th = np.linspace(0.0, 3000.0, 301)
    phi4 = th.copy()
    for _ in range(4):
        phi4 = phi4 - _loss(k, T_ref, phi4)
    got = _poly(c[2:], th) + 2.0 * (c[0] + c[1]) * th
    assert np.abs(got - phi4).max() <= 1e-5
    a1 = _loss(k, T_ref, 1e-3) / 1e-3                # loss'(0), to 1e-9 relative
    assert c[0] == pytest.approx(k.alpha * k.dt * k.inv_dx2 * (1 - a1) ** 3, rel=1e-8)


def _step(th, cx, cy, p):
    u = np.pad(th, 1, mode="edge")                   # mirror edges
    return (_poly(p, th) + cy * (u[:-2, 1:-1] + u[2:, 1:-1])) + cx * (u[1:-1, :-2] + u[1:-1, 2:])


def test_coarse_step_tracks_the_fine_steps_on_a_cooling_ridge():
    k = _fine_const(radiation=True)
    T_ref = 300.0
    cx, cy = k.alpha * k.dt * k.inv_dx2, k.alpha * k.dt * k.inv_dy2
    th_ = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    L = np.polyfit(th_ * 1000.0, _loss(k, T_ref, th_ * 1000.0), 4)[::-1]   # exact quartic
    kc = 1.0 - 2.0 * (cx + cy)
    pf = np.array([-L[0], kc - L[1], -L[2], -L[3], -L[4]])
    ny, nx = 128, 64
    y = np.arange(ny) + 0.5 - ny / 2
    x = np.arange(nx) + 0.5
    th0 = 2000.0 * np.exp(-y[:, None] ** 2 / 50.0) * np.exp(-x[None, :] / 200.0)
    f = th0.copy()
    n = 400                                          # 0.48 ms of cooling
    for _ in range(n):
        f = _step(f, cx, cy, pf)
    R = lambda a: 0.25 * ((a[0::2, 0::2] + a[0::2, 1::2]) + (a[1::2, 0::2] + a[1::2, 1::2]))  # noqa: E731
    want = R(f)
    c = _coeffs(k, T_ref)
    g = R(th0)
    for _ in range(n // 4):
        g = _step(g, c[0], c[1], c[2:])
    plain = R(th0)
    p4 = np.array([-4 * L[0], kc - 4 * L[1], -4 * L[2], -4 * L[3], -4 * L[4]])
    for _ in range(n // 4):
        plain = _step(plain, cx, cy, p4)
    e_new, e_plain = np.abs(g - want).max(), np.abs(plain - want).max()
    assert e_new <= 0.2 and e_plain >= 1.0, (e_new, e_plain)
