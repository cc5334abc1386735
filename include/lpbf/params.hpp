// SPDX-License-Identifier: 0BSD
// Host-side physics and configuration types of the step library.
//
// Same names, fields, defaults and formulas as the reference
// include/params.hpp:9-132 (Material, Process, Domain, dt_cfl, the IN718
// Mills tables and dt_cfl_ext) and include/lpbf_step_const.hpp:36-54
// (make_const), so reference host code compiles against this header
// unchanged.  Everything here is plain host arithmetic in FP64.
#pragma once

#include <algorithm>
#include <cmath>

#include "../lpbf_step_const.h"

namespace lpbf {

enum class Physics { Baseline, Extended };

inline constexpr double SIGMA_SB = 5.670374419e-8;  // Stefan-Boltzmann [W/m^2/K^4]

struct Material {
    double k = 25.0;      // conductivity [W/m/K]
    double rho = 8190.0;  // density [kg/m^3]
    double cp = 565.0;    // heat capacity [J/kg/K]
    double Tm = 1609.0;   // melting point used for the melt-pool metrics [K]
    double eps = 0.40;    // emissivity
    double eta = 0.35;    // absorptivity
    constexpr double alpha() const noexcept { return k / (rho * cp); }
};

struct Process {
    double P = 285.0;        // laser power [W]
    double V = 0.96;         // scan speed [m/s]
    double r0 = 33.5e-6;     // 1/e^2 beam radius [m]
    double h = 40.0e-6;      // z-averaging thickness [m]
    double hConv = 100.0;    // convection coefficient [W/m^2/K]
    double dSub = 30.0e-6;   // substrate conduction length [m]
    double Tamb = 298.15;    // ambient [K]
    double Tsub = 298.15;    // substrate [K]
    bool radiation = true;
};

struct Domain {
    double Lx = 1.6e-3, Ly = 0.6e-3;
    int nx = 240, ny = 120;
    constexpr double dx() const noexcept { return Lx / (nx - 1); }
    constexpr double dy() const noexcept { return Ly / (ny - 1); }
};

static_assert(sizeof(Material) == 48 && sizeof(Process) == 72 && sizeof(Domain) == 24,
              "reference layouts");

// Explicit-Euler diffusion limit with a 0.35 safety factor.
inline double dt_cfl(Material const& m, Domain const& d) noexcept {
    const double a = m.alpha();
    return 0.35 / (a * (1.0 / (d.dx() * d.dx()) + 1.0 / (d.dy() * d.dy())));
}

// IN718 temperature-dependent properties (Mills 2002 piecewise-linear fit,
// latent heat folded into an apparent cp over the mushy interval).
namespace in718 {

inline constexpr double T_SOL = 1533.0;
inline constexpr double T_LIQ = 1609.0;
inline constexpr double LF = 210.0e3;
inline constexpr int N_PTS = 9;
inline constexpr double T_PTS[N_PTS] = {298, 473, 673, 873, 1073, 1273, 1473, 1533, 1609};
inline constexpr double K_PTS[N_PTS] = {8.9, 12.9, 15.8, 18.7, 21.9, 25.6, 28.6, 29.3, 29.6};
inline constexpr double CP_PTS[N_PTS] = {435, 479, 515, 558, 580, 628, 670, 685, 720};

inline double interp(double T, double const* xs, double const* ys, int n) noexcept {
    if (T <= xs[0]) return ys[0];
    if (T >= xs[n - 1]) return ys[n - 1];
    int i = 1;
    while (T > xs[i]) ++i;
    const double f = (T - xs[i - 1]) / (xs[i] - xs[i - 1]);
    return ys[i - 1] + f * (ys[i] - ys[i - 1]);
}

inline double k_of(double T) noexcept { return interp(T, T_PTS, K_PTS, N_PTS); }
inline double cp_of(double T) noexcept { return interp(T, T_PTS, CP_PTS, N_PTS); }

// Melt-pool conductivity multiplier ramping 1 -> kmult across the mushy zone.
inline double k_enh(double T, double kmult) noexcept {
    if (kmult <= 1.0) return 1.0;
    const double s = (T <= T_SOL) ? 0.0 : (T >= T_LIQ ? 1.0 : (T - T_SOL) / (T_LIQ - T_SOL));
    return 1.0 + (kmult - 1.0) * s;
}
inline double k_eff(double T, double kmult) noexcept { return k_of(T) * k_enh(T, kmult); }

inline double cp_eff(double T) noexcept {
    double c = cp_of(T);
    if (T >= T_SOL && T <= T_LIQ) c += LF / (T_LIQ - T_SOL);
    return c;
}

// Largest k_eff/(rho*cp) over T = 298, 303, ..., 2198 K.
inline double alpha_eff_max(double rho, double kmult) noexcept {
    double m = 0.0;
    for (double T = 298.0; T <= 2200.0; T += 5.0)
        m = std::max(m, k_eff(T, kmult) / (rho * cp_of(T)));
    return m;
}

}  // namespace in718

inline double dt_cfl_ext(Material const& m, Domain const& d, double kmult = 1.0) noexcept {
    const double a = in718::alpha_eff_max(m.rho, kmult);
    return 0.35 / (a * (1.0 / (d.dx() * d.dx()) + 1.0 / (d.dy() * d.dy())));
}

}  // namespace lpbf

// Every scalar of one step, from the configuration (host arithmetic).
inline StepConst make_const(lpbf::Material const& mat, lpbf::Process const& pr,
                            lpbf::Domain const& dom, double dt) {
    StepConst k{};
    k.nx = dom.nx;
    k.ny = dom.ny;
    k.dx = dom.dx();
    k.dy = dom.dy();
    k.inv_dx2 = 1.0 / (k.dx * k.dx);
    k.inv_dy2 = 1.0 / (k.dy * k.dy);
    k.inv_r0_sq = 1.0 / (pr.r0 * pr.r0);
    k.alpha = mat.alpha();
    k.K = mat.k;
    k.rho = mat.rho;
    k.cp = mat.cp;
    k.rho_cp = mat.rho * mat.cp;
    k.inv_h = 1.0 / pr.h;
    k.h = pr.h;
    k.dt = dt;
    k.peak_S = (2.0 * mat.eta * pr.P) / (M_PI * pr.r0 * pr.r0);
    k.eta = mat.eta;
    k.P = pr.P;
    k.hConv = pr.hConv;
    k.Tamb = pr.Tamb;
    k.Tsub = pr.Tsub;
    k.dSub = pr.dSub;
    k.eps = mat.eps;
    k.radiation = pr.radiation;
    k.kmult = 1.0;
    k.q_scale = 1.0;
    k.x_l = 0.0;
    k.y_l = 0.0;
    return k;
}
