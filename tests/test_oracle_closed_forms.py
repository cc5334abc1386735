"""Pin the CPU oracle on what the mathematics fixes.

Each check is chosen so that a plausible slip in the oracle (a dropped loss
term, a wrong sign, a swapped dx/dy, an off-by-one at an edge, a wrong tie
rule) fails it: closed-form first step, discrete eigenmode decay under both
boundary conditions, energy identities, tiling independence, reductions
against numpy, and bit equality with the reference's own headers.
"""
import ctypes as C
import math

import numpy as np
import pytest


def _setup(ora, nx=48, ny=36, case="0"):
    s = ora.defaults(case)
    s.dom.nx, s.dom.ny = nx, ny
    return s


# ---------------------------------------------------------------------------
# closed-form first step: T1 = Tamb + dt*S(x, y) on a uniform ambient field
# ---------------------------------------------------------------------------

def test_first_step_closed_form_case0(ora):
    s = ora.defaults("0")
    dt = ora.dt_cfl(s)
    k = ora.make_const(s, dt)
    k.x_l, k.y_l = 0.2 * s.dom.Lx, 0.5 * s.dom.Ly
    T0 = np.full((s.dom.ny, s.dom.nx), s.proc.Tamb)
    T1 = ora.step(T0, k)
    x = np.arange(s.dom.nx) * (s.dom.Lx / (s.dom.nx - 1))
    y = np.arange(s.dom.ny) * (s.dom.Ly / (s.dom.ny - 1))
    r2 = (x[None, :] - k.x_l) ** 2 + (y[:, None] - k.y_l) ** 2
    peak_S = 2 * 0.35 * 285.0 / (math.pi * 33.5e-6 ** 2)
    S = peak_S * np.exp(-2 * r2 / 33.5e-6 ** 2) / (8190.0 * 565.0 * 40e-6)
    np.testing.assert_allclose(T1, s.proc.Tamb + dt * S, rtol=0, atol=1e-10)
    assert abs(T1.max() - 614.770450) < 5e-7


def test_first_step_closed_form_extended(ora):
    s = ora.defaults("0")
    dt = ora.dt_cfl_ext(s)
    k = ora.make_const(s, dt)
    k.x_l, k.y_l = 0.5 * s.dom.Lx, 0.5 * s.dom.Ly
    T0 = np.full((s.dom.ny, s.dom.nx), s.proc.Tamb)
    T1 = ora.step(T0, k, phys=1)
    x = np.arange(s.dom.nx) * k.dx
    y = np.arange(s.dom.ny) * k.dy
    r2 = (x[None, :] - k.x_l) ** 2 + (y[:, None] - k.y_l) ** 2
    q_in = k.peak_S * np.exp(-2 * r2 * k.inv_r0_sq)
    # Mills cp at 298.15 K: first segment of the table
    cp = 435 + (298.15 - 298) / (473 - 298) * (479 - 435)
    np.testing.assert_allclose(T1, s.proc.Tamb + dt * q_in / 40e-6 / (8190.0 * cp),
                               rtol=1e-13, atol=0)


# ---------------------------------------------------------------------------
# discrete eigenmodes: source off, radiation off, Tamb = Tsub
# ---------------------------------------------------------------------------

def _linear_setup(ora, nx, ny, dx, dy):
    s = ora.defaults("0")
    s.dom.nx, s.dom.ny = nx, ny
    s.dom.Lx, s.dom.Ly = dx * (nx - 1), dy * (ny - 1)
    s.proc.radiation = False
    s.proc.Tamb = s.proc.Tsub = 300.0
    s.proc.hConv = 150.0
    dt = 0.8 * ora.dt_cfl(s)
    k = ora.make_const(s, dt)
    k.q_scale = 0.0
    return s, k


@pytest.mark.parametrize("kx,ly", [(1, 0), (0, 2), (3, 1), (7, 5)])
def test_mirror_eigenmode_decay(ora, kx, ly):
    nx, ny, dx, dy = 40, 30, 7e-6, 5e-6
    s, k = _linear_setup(ora, nx, ny, dx, dy)
    X = np.cos(math.pi * kx * (np.arange(nx) + 0.5) / nx)
    Y = np.cos(math.pi * ly * (np.arange(ny) + 0.5) / ny)
    mode = Y[:, None] * X[None, :]
    T = 300.0 + 10.0 * mode
    lam = (k.alpha * (4 / dx ** 2 * math.sin(math.pi * kx / (2 * nx)) ** 2
                      + 4 / dy ** 2 * math.sin(math.pi * ly / (2 * ny)) ** 2)
           + (k.hConv + k.K / k.dSub) / (k.rho * k.cp * k.h))
    g = 1.0 - k.dt * lam
    n = 50
    for _ in range(n):
        T = ora.step(T, k)
    np.testing.assert_allclose(T - 300.0, 10.0 * g ** n * mode, atol=1e-9)


@pytest.mark.parametrize("kx,ly", [(1, 1), (2, 3), (5, 1)])
def test_dirichlet_sine_mode_decay(ora, kx, ly):
    nx, ny, dx, dy = 34, 26, 6e-6, 6e-6
    s, k = _linear_setup(ora, nx, ny, dx, dy)
    X = np.sin(math.pi * kx * np.arange(nx) / (nx - 1))
    Y = np.sin(math.pi * ly * np.arange(ny) / (ny - 1))
    mode = Y[:, None] * X[None, :]
    T = 300.0 + 10.0 * mode
    T[0, :] = T[-1, :] = T[:, 0] = T[:, -1] = 300.0
    lam = (k.alpha * (4 / dx ** 2 * math.sin(math.pi * kx / (2 * (nx - 1))) ** 2
                      + 4 / dy ** 2 * math.sin(math.pi * ly / (2 * (ny - 1))) ** 2)
           + (k.hConv + k.K / k.dSub) / (k.rho * k.cp * k.h))
    g = 1.0 - k.dt * lam
    n = 40
    for _ in range(n):
        T, st = ora.step_ex(T, k, bc=1, T_bc=300.0)
    np.testing.assert_allclose(T - 300.0, 10.0 * g ** n * mode, atol=1e-9)
    assert np.all(T[0, :] == 300.0) and np.all(T[:, -1] == 300.0)


def test_step_ex_mirror_is_bitwise_step(ora):
    s = _setup(ora, 50, 40)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 20 * k.dx, 17 * k.dy
    rng = np.random.default_rng(3)
    T = 300 + 2000 * rng.random((40, 50))
    for phys in (0, 1):
        a = ora.step(T, k, phys)
        b, _ = ora.step_ex(T, k, phys, bc=0)
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# energy identities
# ---------------------------------------------------------------------------

def test_energy_identity_mirror_baseline(ora):
    s = _setup(ora, 64, 48)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 30 * k.dx, 20 * k.dy
    rng = np.random.default_rng(1)
    T = 300 + 1500 * rng.random((48, 64))
    T1, st = ora.step_ex(T, k)
    # sum of the 5-point Laplacian with ghost = centre telescopes to zero
    assert abs(st.sum_dT - (st.sum_src - st.sum_loss)) < 1e-9 * (abs(st.sum_src) + abs(st.sum_loss) + 1)
    assert abs(st.sum_dT - (T1 - T).sum()) < 1e-9 * abs(st.sum_dT)
    assert st.molten == int((T1 > 1609.0).sum())


def test_energy_identity_mirror_extended(ora):
    """With face conductivities shared by both cells the divergence sums to
    zero, so sum(rho*cp_eff*(T'-T)) = dt*sum(q_in - losses)/h."""
    s = _setup(ora, 40, 30)
    k = ora.make_const(s, ora.dt_cfl_ext(s, 2.0))
    k.kmult = 2.0
    k.x_l, k.y_l = 20 * k.dx, 15 * k.dy
    rng = np.random.default_rng(2)
    T = 300 + 2000 * rng.random((30, 40))
    T1 = ora.step(T, k, phys=1)
    cpe = np.vectorize(ora.lib().ora_cp_eff)(T)
    lhs = (k.rho * cpe * (T1 - T)).sum()
    x = np.arange(40) * k.dx
    y = np.arange(30) * k.dy
    r2 = (x[None, :] - k.x_l) ** 2 + (y[:, None] - k.y_l) ** 2
    q_in = k.peak_S * np.exp(-2 * r2 * k.inv_r0_sq)
    q_top = k.hConv * (T - k.Tamb) + k.eps * 5.670374419e-8 * (T ** 4 - k.Tamb ** 4)
    kof = np.vectorize(ora.lib().ora_k_of)(T)
    q_bot = kof / k.dSub * (T - k.Tsub)
    energy = k.dt * ((q_in - q_top - q_bot) / k.h).sum()
    assert abs(lhs - energy) < 1e-9 * np.abs(k.dt * q_in / k.h).sum()
    # the oracle's per-cell source/loss split sums to the same balance
    _, st = ora.step_ex(T, k, phys=1)
    assert abs((T1 - T).sum() - st.sum_dT) < 1e-9 * abs(st.sum_dT)


def test_extended_energy_split(ora):
    """The extended statistics' source and loss sums, pinned from the field.

    Per cell T' = T + dt*(div + S - loss)/(rho*cp_eff(T)*...) with cp_eff,
    div and loss depending on the old field only, so for one field:
      * the source sum is linear in q_scale and equals the change of sum(T'-T)
        when the beam is switched on;
      * the loss sums of two loss models differ by minus the change of
        sum(T'-T) (convection + radiation doubled, then the substrate
        conductance doubled): a dropped or mis-weighted loss term fails;
      * at T = Tamb = Tsub (uniform) every loss vanishes and div = 0, so
        sum_loss = 0 and sum_dT = sum_src.
    The cells under the beam and a quarter of the rest lie in the latent
    window (1533-1609 K), where cp_eff carries the latent heat, so a source
    or loss sum weighted by cp instead of cp_eff fails."""
    s = _setup(ora, 40, 30)
    k = ora.make_const(s, ora.dt_cfl_ext(s, 1.5))
    k.kmult = 1.5
    k.x_l, k.y_l = 17.3 * k.dx, 12.6 * k.dy
    rng = np.random.default_rng(5)
    T = 300 + 2000 * rng.random((30, 40))
    latent = 1540.0 + 60.0 * rng.random((30, 40))
    T = np.where(rng.random((30, 40)) < 0.25, latent, T)
    T[8:18, 12:24] = latent[8:18, 12:24]            # under the beam
    scale = np.abs(T).sum() * 1e-13

    def run(**kw):
        kk = ora.make_const(s, k.dt)
        C.pointer(kk)[0] = k
        for name, v in kw.items():
            setattr(kk, name, v)
        T1, st = ora.step_ex(T, kk, phys=1)
        assert abs((T1 - T).sum() - st.sum_dT) <= scale
        return st

    on, off, two = run(), run(q_scale=0.0), run(q_scale=2.0)
    assert off.sum_src == 0.0 and on.sum_src > 0.0
    assert abs(on.sum_src - (on.sum_dT - off.sum_dT)) <= scale
    assert abs(two.sum_src - 2.0 * on.sum_src) <= scale
    assert on.sum_loss == off.sum_loss == two.sum_loss
    more_top = run(hConv=2.0 * k.hConv, eps=2.0 * k.eps)
    more_bot = run(dSub=0.5 * k.dSub)
    for st in (more_top, more_bot):
        assert st.sum_loss > on.sum_loss
        assert abs((st.sum_loss - on.sum_loss) + (st.sum_dT - on.sum_dT)) <= scale
    # uniform field at the ambient and substrate temperature: no loss, no diffusion
    Tu = np.full((30, 40), 700.0)
    ku = ora.make_const(s, k.dt)
    C.pointer(ku)[0] = k
    ku.Tamb = ku.Tsub = 700.0
    T1, st = ora.step_ex(Tu, ku, phys=1)
    assert st.sum_loss == 0.0
    assert abs(st.sum_dT - st.sum_src) <= 1e-12 * st.sum_src


def test_source_normalisation(ora):
    """sum(peak_S * exp(-2 r^2/r0^2)) dx dy = eta * P for a resolved Gaussian."""
    s = _setup(ora, 200, 160)
    s.dom.Lx, s.dom.Ly = 199 * 5e-6, 159 * 5e-6
    s.proc.Tamb = s.proc.Tsub = 300.0
    k = ora.make_const(s, 1e-7)
    k.x_l, k.y_l = 100 * 5e-6 + 1.3e-6, 80 * 5e-6 - 0.7e-6
    T = np.full((160, 200), 300.0)
    _, st = ora.step_ex(T, k)
    e_in = st.sum_src * k.rho_cp * k.h * k.dx * k.dy / k.dt
    assert abs(e_in - k.eta * k.P) < 1e-9 * k.eta * k.P


# ---------------------------------------------------------------------------
# tiling, reductions, metrics
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ntiles", [1, 2, 3, 7, 36])
def test_tiling_independence(ora, ntiles):
    s = _setup(ora, 30, 36)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 11 * k.dx, 0.0
    rng = np.random.default_rng(ntiles)
    T = 300 + 1000 * rng.random((36, 30))
    for phys in (0, 1):
        assert np.array_equal(ora.step_tiled(T, k, phys, ntiles), ora.step(T, k, phys))


def test_mirror_symmetry_about_laser_row(ora):
    """Source centred between rows j and j+1 of a mirror-BC grid with an even
    row count: the field stays symmetric (y -> ny-1-y) at every step."""
    s = _setup(ora, 40, 24)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 13 * k.dx, 11.5 * k.dy
    T = np.full((24, 40), s.proc.Tamb)
    for _ in range(30):
        T = ora.step(T, k)
    np.testing.assert_allclose(T, T[::-1, :], rtol=1e-13, atol=0)


def test_linear_in_power_without_radiation(ora):
    s = _setup(ora, 40, 30)
    s.proc.radiation = False
    s.proc.Tamb = s.proc.Tsub = 300.0
    dt = ora.dt_cfl(s)
    out = []
    for P in (100.0, 300.0):
        s.proc.P = P
        k = ora.make_const(s, dt)
        k.x_l, k.y_l = 15 * k.dx, 14.2 * k.dy
        T = np.full((30, 40), 300.0)
        for _ in range(20):
            T = ora.step(T, k)
        out.append(T - 300.0)
    np.testing.assert_allclose(out[1], 3.0 * out[0], rtol=1e-11, atol=1e-12)


def test_reduce_matches_numpy_with_ties(ora):
    nx, ny, Tm = 37, 23, 1609.0
    rng = np.random.default_rng(5)
    T = 300 + 1000 * rng.random((ny, nx))
    T[3, 30] = T[12, 4] = T[20, 36] = 5000.0   # three-way tie at the max
    T[15, 2] = 1700.0
    T[5, 10] = 1610.0
    for nblocks in (1, 2, 3, 4):
        parts = ora.reduce_partials(T, Tm, nblocks)
        g = ora.reduce_finalize(parts, nx, ny)
        assert g.tmax == 5000.0
        assert g.arg == 3 * nx + 30           # smallest row-major index
        rows = np.nonzero((T > Tm).any(axis=1))[0]
        assert (g.ymin, g.ymax) == (rows.min(), rows.max())


def test_reduce_identity_when_nothing_molten(ora):
    T = np.full((5, 7), 300.0)
    g = ora.reduce_finalize(ora.reduce_partials(T, 1609.0, 2), 7, 5)
    assert (g.ymin, g.ymax) == (5, -1) and g.arg == 0 and g.tmax == 300.0


def test_measure_against_numpy(ora):
    ny, nx, dy, Tm = 30, 25, 4e-6, 1609.0
    y = np.arange(ny)
    x = np.arange(nx)
    T = 300 + 3000 * np.exp(-((y[:, None] - 14.3) ** 2 / 20 + (x[None, :] - 9) ** 2 / 30))
    m = ora.measure(T, dy, Tm)
    assert m.peak_T == T.max() and m.x_hot == int(np.argmax(T) % nx)
    rows = np.nonzero((T > Tm).any(axis=1))[0]
    assert m.W_um == pytest.approx((rows.max() - rows.min() + 1) * dy * 1e6, rel=1e-15)
    col = T[:, m.x_hot]
    a, b = np.nonzero(col > Tm)[0][[0, -1]]
    ylo = (a - 1 + (Tm - col[a - 1]) / (col[a] - col[a - 1])) * dy
    yhi = (b + (Tm - col[b]) / (col[b + 1] - col[b])) * dy
    assert m.W_sub_um == pytest.approx((yhi - ylo) * 1e6, rel=1e-14)
    assert ora.wsub_from_col(col, dy, Tm) == m.W_sub_um


def test_step_matches_numpy_brute_force(ora):
    """Independent array formulation of the baseline update on a tiny grid."""
    s = _setup(ora, 9, 7)
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 3.3 * k.dx, 2.1 * k.dy
    rng = np.random.default_rng(11)
    T = 300 + 2500 * rng.random((7, 9))
    P = np.pad(T, 1, mode="edge")  # ghost = centre
    lap = ((P[1:-1, 2:] + P[1:-1, :-2] - 2 * T) * k.inv_dx2
           + (P[2:, 1:-1] + P[:-2, 1:-1] - 2 * T) * k.inv_dy2)
    x = np.arange(9) * k.dx
    y = np.arange(7) * k.dy
    r2 = (x[None, :] - k.x_l) ** 2 + (y[:, None] - k.y_l) ** 2
    S = k.peak_S * np.exp(-2 * r2 * k.inv_r0_sq) / (k.rho_cp * k.h)
    loss = (k.hConv * (T - k.Tamb) + k.eps * 5.670374419e-8 * (T ** 4 - k.Tamb ** 4)
            + k.K / k.dSub * (T - k.Tsub)) / k.h / k.rho_cp
    ref = T + k.dt * (k.alpha * lap + S - loss)
    np.testing.assert_allclose(ora.step(T, k), ref, rtol=1e-14, atol=1e-11)


# ---------------------------------------------------------------------------
# IN718 tables and the reference headers
# ---------------------------------------------------------------------------

KN_T = [298, 473, 673, 873, 1073, 1273, 1473, 1533, 1609]
KN_K = [8.9, 12.9, 15.8, 18.7, 21.9, 25.6, 28.6, 29.3, 29.6]
KN_CP = [435, 479, 515, 558, 580, 628, 670, 685, 720]


def test_tables_at_knots_and_midpoints(ora):
    L = ora.lib()
    for i, T in enumerate(KN_T):
        assert L.ora_k_of(T) == KN_K[i] and L.ora_cp_of(T) == KN_CP[i]
    for i in range(8):
        Tm_ = 0.5 * (KN_T[i] + KN_T[i + 1])
        assert L.ora_k_of(Tm_) == pytest.approx(0.5 * (KN_K[i] + KN_K[i + 1]), rel=1e-15)
    assert L.ora_k_of(100.0) == 8.9 and L.ora_k_of(5000.0) == 29.6
    assert L.ora_k_enh(2000.0, 1.0) == 1.0 and L.ora_k_enh(2000.0, 3.0) == 3.0
    assert L.ora_k_enh(1571.0, 3.0) == pytest.approx(2.0, rel=1e-15)
    assert L.ora_cp_eff(1600.0) - L.ora_cp_of(1600.0) == pytest.approx(210e3 / 76, rel=1e-15)
    assert L.ora_cp_eff(1700.0) == 720.0


def test_oracle_equals_reference_headers(ora):
    ref = ora.ref_probe()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent when building)")
    L = ora.lib()
    assert ref.ref_sizeof_stepconst() == 216 and ref.ref_sizeof_partial() == 24
    for T in np.linspace(200, 2500, 461):
        for km in (1.0, 1.7, 3.0):
            assert L.ora_k_eff(T, km) == ref.ref_k_eff(T, km)
        assert L.ora_cp_eff(T) == ref.ref_cp_eff(T)
    for case in ("0", "1.1", "1.2", "2.1", "2.2", "3.1", "3.2"):
        s = ora.defaults(case)
        mat = np.array([s.mat.k, s.mat.rho, s.mat.cp, s.mat.Tm, s.mat.eps, s.mat.eta])
        proc = np.array([s.proc.P, s.proc.V, s.proc.r0, s.proc.h, s.proc.hConv,
                         s.proc.dSub, s.proc.Tamb, s.proc.Tsub])
        dp = C.POINTER(C.c_double)
        args = (mat.ctypes.data_as(dp), s.dom.Lx, s.dom.Ly, s.dom.nx, s.dom.ny)
        assert ora.dt_cfl(s) == ref.ref_dt_cfl(*args)
        for km in (1.0, 2.5):
            assert ora.dt_cfl_ext(s, km) == ref.ref_dt_cfl_ext(*args, km)
        dt = ora.dt_cfl(s)
        a = ora.make_const(s, dt)
        b = ora.StepConst()
        ref.ref_make_const(mat.ctypes.data_as(dp), proc.ctypes.data_as(dp), 1,
                           s.dom.Lx, s.dom.Ly, s.dom.nx, s.dom.ny, dt, C.byref(b))
        for name, _ in ora.StepConst._fields_:
            assert getattr(a, name) == getattr(b, name), name


# ---------------------------------------------------------------------------
# new inputs: noise and raster
# ---------------------------------------------------------------------------

def test_noise_known_answer_and_range(ora):
    # SplitMix64's first output from state 0 is 0xE220A8397B1DCDAF (published)
    assert ora.noise(0, 0) == (0xE220A8397B1DCDAF >> 40) / 2 ** 24 - 0.5
    v = ora.noise_field(7, 300.0, 0.5, (64, 64)) - 300.0
    assert v.min() >= -0.5 and v.max() < 0.5
    assert abs(v.mean()) < 0.02 and abs(v.std() - 1 / math.sqrt(12)) < 0.02
    assert np.all(v * 2 ** 24 == np.round(v * 2 ** 24))   # 24-bit exact
    w = ora.noise_field(8, 300.0, 0.5, (64, 64)) - 300.0
    assert not np.array_equal(v, w)


def test_raster_positions(ora):
    r = ora.OraRaster(x0=1e-3, y0=2e-3, length=4e-4, hatch=8e-5, n_tracks=3, V=1.0)
    dt = 1e-5                      # 10 um per step; 40 steps per track, 8 per jump
    p = ora.raster_plan(r, dt, 200)
    assert tuple(p[0]) == (1e-3, 2e-3, 1.0)
    np.testing.assert_allclose(p[10], [1e-3 + 1e-4, 2e-3, 1.0], rtol=1e-12)
    # track 0 ends at step 40 (d = 400 um): beam off, moving up the hatch
    assert p[41, 2] == 0.0 and p[41, 0] == pytest.approx(1.4e-3, rel=1e-12)
    # track 1 runs in -x starting at step 48
    np.testing.assert_allclose(p[50, :2], [1.4e-3 - 2e-5, 2e-3 + 8e-5], rtol=1e-9)
    assert p[50, 2] == 1.0
    # after the last track the beam stays off at the end point
    assert np.all(p[140:, 2] == 0.0)
    np.testing.assert_allclose(p[199, :2], [1.4e-3, 2e-3 + 1.6e-4], rtol=1e-12)


def test_threaded_rows_equal_single_step(ora):
    """The threaded timing path of the oracle computes exactly ora_step."""
    from concurrent.futures import ThreadPoolExecutor
    s = ora.defaults("0")
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 0.5e-3, 0.3e-3
    rng = np.random.default_rng(3)
    T = 300.0 + 2000.0 * rng.random((s.dom.ny, s.dom.nx))
    with ThreadPoolExecutor(5) as pool:
        got = ora.step_threads(T, k, 0, threads=5, pool=pool)
    assert np.array_equal(got, ora.step(T, k, 0))


def test_fp32_cpu_baseline_step_within_fp32_rounding(ora):
    """The FP32 CPU baseline step (timed beside the FP32 GPU path) against the
    FP64 oracle, on a field where diffusion, source, convection, radiation and
    substrate loss each move a cell by far more than FP32 rounding (edges
    included), so a dropped or wrong term fails the bound."""
    s = _setup(ora, 64, 48)
    s.proc.hConv = 1e5
    s.mat.eps = 1.0
    k = ora.make_const(s, ora.dt_cfl(s))
    k.x_l, k.y_l = 31.4 * k.dx, 22.7 * k.dy
    rng = np.random.default_rng(11)
    T = 300 + 2000 * rng.random((48, 64))
    ref = ora.step(T, k)
    got = ora.step_threads_f32(np.ascontiguousarray(T, dtype=np.float32), k, threads=3)
    # FP32 rounding of values up to ~3000 K over a dozen operations
    assert np.abs(got - ref).max() <= 5e-3
    # every cell moves by 100x the bound or more (the planted slips checked on
    # the CPU: radiation, substrate loss or source dropped, lap_y scaled
    # by 1.001, each fail the first bound)
    assert np.abs(ref - T).min() > 0.5


@pytest.mark.parametrize("ntiles,nthreads", [(1, 1), (6, 3), (16, 8), (40, 5)])
def test_steptask_run_equals_full_grid_steps(ora, ntiles, nthreads):
    """The CPU baseline's step-task loop (tiles built, stepped by a thread
    pool, stitched in order) gives exactly the full-grid oracle steps, in
    FP64 (both physics) and FP32, for any tile and thread count (tiling
    independence, lpbf_cuda.cu:14-18)."""
    s = ora.defaults("0")
    s.dom.nx, s.dom.ny = 97, 37
    k = ora.make_const(s, ora.dt_cfl(s))
    n = 7
    plan = ora.track_plan(3.0, k.dt, s.dom.Lx, s.dom.Ly, n)
    rng = np.random.default_rng(ntiles)
    T0 = 300.0 + 1500.0 * rng.random((37, 97))
    for phys in (0, 1):
        want = T0.copy()
        for q in range(n):
            kk = ora.StepConst.from_buffer_copy(bytes(k))
            kk.x_l, kk.y_l, kk.q_scale = plan[q]
            want = ora.step(want, kk, phys)
        got = T0.copy()
        ph = ora.steptask_run(got, k, plan, n, ntiles, nthreads, phys)
        assert np.array_equal(got, want), phys
        assert ph[2] >= ph[0] >= 0.0 and ph[2] >= ph[1] >= 0.0
    want = np.ascontiguousarray(T0, dtype=np.float32)
    for q in range(n):
        kk = ora.StepConst.from_buffer_copy(bytes(k))
        kk.x_l, kk.y_l, kk.q_scale = plan[q]
        want = ora.step_threads_f32(want, kk, threads=1)
    got = np.ascontiguousarray(T0, dtype=np.float32)
    ora.steptask_run(got, k, plan, n, ntiles, nthreads)
    assert np.array_equal(got, want)
    with pytest.raises(RuntimeError):
        ora.steptask_run(np.ascontiguousarray(T0, dtype=np.float32), k, plan, n, 4, 2, phys=1)
