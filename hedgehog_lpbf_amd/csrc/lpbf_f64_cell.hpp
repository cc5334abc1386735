// SPDX-License-Identifier: 0BSD
// The FP64 cell update shared by the FP64 step kernels: the reference
// expression in the reference's operation order (include/step_task.hpp:93-129
// baseline, :132-187 extended; src/gpu/kernels.cu:106-147), with FP
// contraction off so every product and sum rounds as written.  Divisions by
// launch constants use the correctly rounded reciprocal form (div_rn), which
// returns the IEEE quotient; the only difference from the CPU restatement is
// the device exp().
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "lpbf_step_const.h"

namespace lpbf_amd {
namespace {

constexpr double kSigmaSB = 5.670374419e-8;  // params.hpp:11

// IN718 Mills tables (params.hpp:67-73) and their interpolation (:75-111).
__constant__ double cT_PTS[9] = {298, 473, 673, 873, 1073, 1273, 1473, 1533, 1609};
__constant__ double cK_PTS[9] = {8.9, 12.9, 15.8, 18.7, 21.9, 25.6, 28.6, 29.3, 29.6};
__constant__ double cCP_PTS[9] = {435, 479, 515, 558, 580, 628, 670, 685, 720};
// RN(1 / (T_PTS[i] - T_PTS[i-1])), folded by the compiler (correctly rounded)
__constant__ double cT_RSTEP[9] = {0.0,
                                   1.0 / (473.0 - 298.0),   1.0 / (673.0 - 473.0),
                                   1.0 / (873.0 - 673.0),   1.0 / (1073.0 - 873.0),
                                   1.0 / (1273.0 - 1073.0), 1.0 / (1473.0 - 1273.0),
                                   1.0 / (1533.0 - 1473.0), 1.0 / (1609.0 - 1533.0)};

// x / c, correctly rounded, for a divisor whose reciprocal r = RN(1/c) is
// known: q = RN(x*r) is within one ulp of x/c, so the remainder x - q*c is
// exact (one FMA) and RN(q + r*(x - q*c)) is RN(x/c) (Markstein's theorem;
// no operand here is near overflow or underflow).  Three FP64 operations
// instead of the IEEE division sequence, same result.
__device__ __forceinline__ double div_rn(double x, double c, double r) {
    const double q = x * r;
    const double e = fma(-q, c, x);
    return fma(e, r, q);
}

__device__ inline double interp9(double T, const double* xs, const double* ys) {
    if (T <= xs[0]) return ys[0];
    if (T >= xs[8]) return ys[8];
    int i = 1;
    while (T > xs[i]) ++i;
    const double f = div_rn(T - xs[i - 1], xs[i] - xs[i - 1], cT_RSTEP[i]);
    return ys[i - 1] + f * (ys[i] - ys[i - 1]);
}
__device__ inline double k_of(double T) { return interp9(T, cT_PTS, cK_PTS); }
__device__ inline double cp_of(double T) { return interp9(T, cT_PTS, cCP_PTS); }
__device__ inline double k_enh(double T, double kmult) {
    if (kmult <= 1.0) return 1.0;
    double s;
    if (T <= 1533.0) s = 0.0;
    else if (T >= 1609.0) s = 1.0;
    else s = div_rn(T - 1533.0, 1609.0 - 1533.0, 1.0 / (1609.0 - 1533.0));
    return 1.0 + (kmult - 1.0) * s;
}
__device__ inline double k_eff(double T, double kmult) { return k_of(T) * k_enh(T, kmult); }
__device__ inline double cp_eff(double T) {
    double c = cp_of(T);
    if (T >= 1533.0 && T <= 1609.0) c += 210.0e3 / (1609.0 - 1533.0);
    return c;
}

// exp(x) is +0 for every x below this in both glibc and the device libm, and
// then S = (q*peak*0)/(rho_cp*h) = +0, so the call can be skipped bit-exactly.
constexpr double kExpZero = -746.0;

// Beam position and power scale of one step (StepConst x_l, y_l, q_scale).
struct Beam {
    double x_l, y_l, q;
};
__device__ __forceinline__ Beam beam_of(const StepConst& k) { return Beam{k.x_l, k.y_l, k.q_scale}; }

// Quotients of launch constants (IEEE divisions, so the host and the device
// compute the same bits): three reciprocals of divisors and the substrate
// conductance K / dSub of the baseline loss.
struct Recips {
    double rho_cp;      // 1 / rho_cp
    double rho_cp_h;    // 1 / (rho_cp * h)
    double dSub;        // 1 / dSub
    double K_dSub;      // K / dSub
};
__host__ __device__ __forceinline__ Recips recips(const StepConst& k) {
    return Recips{1.0 / k.rho_cp, 1.0 / (k.rho_cp * k.h), 1.0 / k.dSub, k.K / k.dSub};
}

// True when every cell of the row at ypos has a Gaussian exponent below
// kExpZero: the exponent -2*r2*inv_r0_sq is computed from
// r2 = dx^2 + dy^2 >= dy^2 by monotone rounded operations, so it is at most
// the value formed from dy^2 alone.
__device__ __forceinline__ bool beam_far(double ypos, const Beam& bm, const StepConst& k) {
    const double b2 = (ypos - bm.y_l) * (ypos - bm.y_l);
    return -2.0 * b2 * k.inv_r0_sq < kExpZero;
}

// Baseline update (step_task.hpp:93-129).  xe: -1 at x = 0, +1 at x = nx-1.
__device__ __forceinline__ double cell_base(double T, double Txm, double Txp,
                                            double Tym, double Typ, int xe,
                                            double xpos, double ypos, bool far,
                                            const Beam& bm, const StepConst& k,
                                            const Recips& rc, double& src, double& lossv) {
    double lap_x;
    if (xe < 0) lap_x = (Txp - T) * k.inv_dx2;
    else if (xe > 0) lap_x = (Txm - T) * k.inv_dx2;
    else lap_x = (Txp - 2.0 * T + Txm) * k.inv_dx2;
    const double lap_y = (Typ - 2.0 * T + Tym) * k.inv_dy2;
    double S = 0.0;
    if (!far) {
        const double r2 = (xpos - bm.x_l) * (xpos - bm.x_l) + (ypos - bm.y_l) * (ypos - bm.y_l);
        const double arg = -2.0 * r2 * k.inv_r0_sq;
        if (!(arg < kExpZero))
            S = div_rn(bm.q * k.peak_S * exp(arg), k.rho_cp * k.h, rc.rho_cp_h);
    }
    double q_top = k.hConv * (T - k.Tamb);
    if (k.radiation) {
        const double T4 = T * T * T * T;
        const double Ta4 = k.Tamb * k.Tamb * k.Tamb * k.Tamb;
        q_top += k.eps * kSigmaSB * (T4 - Ta4);
    }
    const double q_bot = rc.K_dSub * (T - k.Tsub);
    const double loss = div_rn((q_top + q_bot) * k.inv_h, k.rho_cp, rc.rho_cp);
    const double dTdt = k.alpha * (lap_x + lap_y) + S - loss;
    src = k.dt * S;
    lossv = k.dt * loss;
    return T + k.dt * dTdt;
}

// Extended update (step_task.hpp:143-185); neighbours already ghost-substituted.
__device__ __forceinline__ double cell_ext(double T, double Txm, double Txp,
                                           double Tym, double Typ, double xpos,
                                           double ypos, bool far, const Beam& bm,
                                           const StepConst& k, const Recips& rc, double& src,
                                           double& lossv) {
    const double km = k.kmult;
    const double kCe = k_eff(T, km);
    const double kxm = 0.5 * (kCe + k_eff(Txm, km));
    const double kxp = 0.5 * (kCe + k_eff(Txp, km));
    const double kym = 0.5 * (kCe + k_eff(Tym, km));
    const double kyp = 0.5 * (kCe + k_eff(Typ, km));
    const double div_k_grad = (kxp * (Txp - T) - kxm * (T - Txm)) * k.inv_dx2
                            + (kyp * (Typ - T) - kym * (T - Tym)) * k.inv_dy2;
    double q_in = 0.0;
    if (!far) {
        const double r2 = (xpos - bm.x_l) * (xpos - bm.x_l) + (ypos - bm.y_l) * (ypos - bm.y_l);
        const double arg = -2.0 * r2 * k.inv_r0_sq;
        if (!(arg < kExpZero)) q_in = bm.q * k.peak_S * exp(arg);
    }
    double q_top = k.hConv * (T - k.Tamb);
    if (k.radiation) {
        const double T4 = T * T * T * T;
        const double Ta4 = k.Tamb * k.Tamb * k.Tamb * k.Tamb;
        q_top += k.eps * kSigmaSB * (T4 - Ta4);
    }
    const double q_bot = div_rn(k_of(T), k.dSub, rc.dSub) * (T - k.Tsub);
    const double rhs = div_k_grad + (q_in - q_top - q_bot) * k.inv_h;
    const double rcp = k.rho * cp_eff(T);
    const double r_rcp = 1.0 / rcp;   // one IEEE division, three quotients
    src = div_rn(k.dt * (q_in * k.inv_h), rcp, r_rcp);
    lossv = div_rn(k.dt * ((q_top + q_bot) * k.inv_h), rcp, r_rcp);
    return T + div_rn(k.dt * rhs, rcp, r_rcp);
}

// Whole-wavefront lane moves of a double without LDS: DPP wave_shr:1 /
// wave_shl:1 on both halves (lane l receives lane l-1 / l+1; the lane
// without a source receives 0), and a lane broadcast through readlane.
__device__ __forceinline__ double lane_from_left(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_right(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_read(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

}  // namespace
}  // namespace lpbf_amd
