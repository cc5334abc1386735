// SPDX-License-Identifier: 0BSD
// FP32 form of the extended physics (include/lpbf_kernels.h,
// lpbf_launch_step_f32_ext): the reference's phys = 1 update
// (include/step_task.hpp:132-187, src/gpu/kernels.cu:125-147) on the FP32
// theta field, one step per launch, with the per-step statistics fused.
//
// Per cell, with T = T_ref + theta:
//   k_eff(T) = k(T) * enh(T), k and cp piecewise linear in the IN718 table
//   (params.hpp:63-73), enh ramping 1 -> kmult over [T_sol, T_liq];
//   face conductivities are the means of the two cells' k_eff;
//   rhs = div(k grad T) + (q_in - q_top - q_bot) / h,
//   T' = T + dt * rhs / (rho * cp_eff(T)), cp_eff = cp + latent heat over
//   the mushy interval.
// The table interpolation is branch-free: for the eight intervals
// [T_i, T_i+1], t_i = med3(T - T_i, 0, T_i+1 - T_i) and
// k(T) = k_0 + sum_i slope_i t_i (the same ramps give cp), which is the
// piecewise-linear interpolation, constant outside the table.  Temperature
// differences are theta differences (exact for the neighbours), and the top
// loss hConv (T - Tamb) + eps sigma (T^4 - Tamb^4) is a polynomial in theta
// expanded around T_ref on the host in FP64, so no large temperatures
// cancel in FP32.
//
// Schedule: a wavefront owns 128 columns (lane l holds columns c0+l and
// c0+64+l, each row one contiguous 256-byte access per half) and a band of
// R rows, streamed top to bottom with the rows above and below in registers
// and the next two rows in flight.  Every row's k_eff is evaluated once, when
// the row is loaded, and moves down the window with it.  Loads clamp their
// column and row into the field: at a mirror edge the clamped neighbour is
// the cell itself, which is the reference's ghost = centre substitution, so
// the mirror condition needs no code; Dirichlet edge cells are held.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lpbf_common.hpp"
#include "lpbf_kernels.h"

namespace lpbf_amd {
namespace {

constexpr int kT = 256;
constexpr int kCols = 128;                   // columns per wavefront
constexpr int kRowsMax = 32;

struct ExtArgs {
    const float* in;
    float* out;
    int nx, ny, ld;
    int R, nwx, ntasks;
    int bc;
    float theta_bc, theta_m;
    float Tref;
    float inv_dx2, inv_dy2, inv_h, dt, rho;
    float tsub_off, inv_dSub;                // T_ref - Tsub, 1 / dSub
    float e0, e1, e2, e3, e4;                // q_top(theta), W/m^2
    float enh_a, enh_b;                      // enh(T) = 1 + clamp(T a + b, 0, 1) (kmult - 1)
    float kmult_m1;                          // kmult - 1 (0: no enhancement)
    float T_sol, T_liq, latent;              // mushy interval and LF / (T_liq - T_sol)
    float rT[8], rW[8], kS[8], cS[8];        // ramps: start, width, k and cp slopes
    float k0, cp0;
    float ax, ay, cut;                       // source in cell units (base-2 exponent)
    double inv_dx, inv_dy, qS0;              // 1/dx, 1/dy, peak_S
    StepStats* part;                         // [nwg]
};

struct Props {
    float k;     // k(T), the table value (substrate loss)
    float ke;    // k_eff(T)
};

__device__ __forceinline__ Props props(const ExtArgs& a, float T) {
    float kv = a.k0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        kv = fmaf(a.kS[i], __builtin_amdgcn_fmed3f(T - a.rT[i], 0.f, a.rW[i]), kv);
    float ke = kv;
    if (a.kmult_m1 != 0.f) {   // launch-uniform
        const float s = __builtin_amdgcn_fmed3f(fmaf(T, a.enh_a, a.enh_b), 0.f, 1.f);
        ke = kv * fmaf(a.kmult_m1, s, 1.f);
    }
    return Props{kv, ke};
}

__device__ __forceinline__ float cp_eff(const ExtArgs& a, float T) {
    float c = a.cp0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        c = fmaf(a.cS[i], __builtin_amdgcn_fmed3f(T - a.rT[i], 0.f, a.rW[i]), c);
    return (T >= a.T_sol && T <= a.T_liq) ? c + a.latent : c;
}

__device__ __forceinline__ float lane_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_at(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

using gfloat = __attribute__((address_space(1))) float;

__global__ __launch_bounds__(kT) void k_step_f32_ext(ExtArgs a, const LaserTick* __restrict__ plan,
                                                     int step) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int task = blockIdx.x * (kT / kWave) + wave;
    const int nx = a.nx, ny = a.ny;

    double tmax = -1.0;
    long long arg = static_cast<long long>(nx) * ny;
    int ymin = ny, ymax = -1;
    long long molten = 0;
    double s_dT = 0.0, s_src = 0.0, s_loss = 0.0;
    float bv = -3.0e38f;
    int bpos = -1;
    int xs0 = 0, xs1 = 0;

    if (task < a.ntasks) {
        const LaserTick t = plan[step];
        const float xc = static_cast<float>(t.x_l * a.inv_dx);
        const float yc = static_cast<float>(t.y_l * a.inv_dy);
        const float qS = static_cast<float>(a.qS0 * t.q_scale);
        const float wcut = sqrtf(a.cut / a.ay);
        int srlo = static_cast<int>(floorf(yc - wcut)) - 1;
        int srhi = static_cast<int>(ceilf(yc + wcut)) + 1;
        if (t.q_scale == 0.0) { srlo = 1; srhi = 0; }

        const int tx = task % a.nwx, ty = task / a.nwx;
        const int c0 = tx * kCols;
        const int ya = ty * a.R, yb = min(ya + a.R, ny);
        const int x0 = c0 + lane, x1 = c0 + kWave + lane;
        xs0 = x0;
        xs1 = x1;
        const int lx0 = min(x0, nx - 1), lx1 = min(x1, nx - 1);
        const int ecol = lane == 0 ? c0 - 1 : c0 + kCols;
        const int lxe = min(max(ecol, 0), nx - 1);
        const gfloat* gin = (const gfloat*)a.in;
        gfloat* gout = (gfloat*)a.out;
        auto row = [&](int r) { return gin + static_cast<long long>(min(max(r, 0), ny - 1)) * a.ld; };

        // the window: rows y-1 (p), y (c), y+1 (n) of both columns, with
        // their k_eff, and the edge column of row y; rows y+2, y+3 in flight
        const gfloat* rp = row(ya - 1);
        const gfloat* rc = row(ya);
        const gfloat* rn = row(ya + 1);
        float p0 = rp[lx0], p1 = rp[lx1];
        float c0v = rc[lx0], c1v = rc[lx1], ce = rc[lxe];
        float n0 = rn[lx0], n1 = rn[lx1], ne = rn[lxe];
        float q0[2], q1[2], qe[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const gfloat* rq = row(ya + 2 + d);
            q0[d] = rq[lx0];
            q1[d] = rq[lx1];
            qe[d] = rq[lxe];
        }
        Props pp0 = props(a, a.Tref + p0), pp1 = props(a, a.Tref + p1);
        Props pc0 = props(a, a.Tref + c0v), pc1 = props(a, a.Tref + c1v);
        Props pn0 = props(a, a.Tref + n0), pn1 = props(a, a.Tref + n1);

        for (int y = ya; y < yb; ++y) {
            const float kce = props(a, a.Tref + ce).ke;
            // x neighbours: lane shifts; lane 0 / 63 take the edge column or
            // the other half's end
            const float up0 = lane_left(c0v), up1 = lane_left(c1v);
            const float dn0 = lane_right(c0v), dn1 = lane_right(c1v);
            const float kup0 = lane_left(pc0.ke), kup1 = lane_left(pc1.ke);
            const float kdn0 = lane_right(pc0.ke), kdn1 = lane_right(pc1.ke);
            const float last0 = lane_at(c0v, kWave - 1), first1 = lane_at(c1v, 0);
            const float klast0 = lane_at(pc0.ke, kWave - 1), kfirst1 = lane_at(pc1.ke, 0);
            const float txm[2] = {lane ? up0 : ce, lane ? up1 : last0};
            const float txp[2] = {lane < kWave - 1 ? dn0 : first1, lane < kWave - 1 ? dn1 : ce};
            const float kxm_n[2] = {lane ? kup0 : kce, lane ? kup1 : klast0};
            const float kxp_n[2] = {lane < kWave - 1 ? kdn0 : kfirst1, lane < kWave - 1 ? kdn1 : kce};
            const float tc[2] = {c0v, c1v}, tp[2] = {p0, p1}, tn[2] = {n0, n1};
            const Props pc[2] = {pc0, pc1};
            const float kp[2] = {pp0.ke, pp1.ke}, kn[2] = {pn0.ke, pn1.ke};
            const bool src_row = y >= srlo && y <= srhi;
            const float ey = a.ay * (static_cast<float>(y) - yc) * (static_cast<float>(y) - yc);
            gfloat* orow = gout + static_cast<long long>(y) * a.ld;
            bool mol[2] = {false, false};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int x = j ? x1 : x0;
                if (x >= nx) continue;
                const float th = tc[j];
                float v, src = 0.f, lossv = 0.f;
                if (a.bc == LPBF_BC_DIRICHLET && (x == 0 || x == nx - 1 || y == 0 || y == ny - 1)) {
                    v = a.theta_bc;
                } else {
                    const float T = a.Tref + th;
                    const float kc = pc[j].ke;
                    // face conductivities, with the mirror ghost (clamped
                    // neighbour = the cell) giving k_face = k_eff(T)
                    const float kxm = 0.5f * (kc + kxm_n[j]), kxp = 0.5f * (kc + kxp_n[j]);
                    const float kym = 0.5f * (kc + kp[j]), kyp = 0.5f * (kc + kn[j]);
                    const float div = (kxp * (txp[j] - th) - kxm * (th - txm[j])) * a.inv_dx2 +
                                      (kyp * (tn[j] - th) - kym * (th - tp[j])) * a.inv_dy2;
                    float qin = 0.f;
                    if (src_row) {
                        const float ux = static_cast<float>(x) - xc;
                        qin = qS * __builtin_amdgcn_exp2f(-fmaf(a.ax * ux, ux, ey));
                    }
                    const float qtop = fmaf(fmaf(fmaf(fmaf(a.e4, th, a.e3), th, a.e2), th, a.e1),
                                            th, a.e0);
                    const float qbot = pc[j].k * a.inv_dSub * (th + a.tsub_off);
                    const float rhs = div + (qin - qtop - qbot) * a.inv_h;
                    const float r = __builtin_amdgcn_rcpf(a.rho * cp_eff(a, T)) * a.dt;
                    v = th + rhs * r;
                    src = qin * a.inv_h * r;
                    lossv = (qtop + qbot) * a.inv_h * r;
                }
                orow[x] = v;
                if (v > bv) {   // a lane meets its cells in increasing index order
                    bv = v;
                    bpos = 2 * y + j;
                }
                mol[j] = v > a.theta_m;
                s_dT += static_cast<double>(v) - static_cast<double>(
                    (a.bc == LPBF_BC_DIRICHLET && (x == 0 || x == nx - 1 || y == 0 || y == ny - 1))
                        ? v : th);
                s_src += src;
                s_loss += lossv;
            }
            const unsigned long long b0 = __ballot(mol[0]), b1 = __ballot(mol[1]);
            molten += __popcll(b0) + __popcll(b1);
            if (b0 | b1) {
                ymin = min(ymin, y);
                ymax = max(ymax, y);
            }
            // slide the window down one row
            p0 = c0v; p1 = c1v; pp0 = pc0; pp1 = pc1;
            c0v = n0; c1v = n1; ce = ne; pc0 = pn0; pc1 = pn1;
            n0 = q0[0]; n1 = q1[0]; ne = qe[0];
            q0[0] = q0[1]; q1[0] = q1[1]; qe[0] = qe[1];
            if (y + 1 < yb) {
                const gfloat* rq = row(y + 4);
                q0[1] = rq[lx0];
                q1[1] = rq[lx1];
                qe[1] = rq[lxe];
                pn0 = props(a, a.Tref + n0);
                pn1 = props(a, a.Tref + n1);
            }
        }
    }

    // statistics: hottest cell (smallest index on ties), sums, molten
    if (bpos >= 0) {
        tmax = static_cast<double>(a.Tref) + static_cast<double>(bv);
        arg = static_cast<long long>(bpos >> 1) * nx + ((bpos & 1) ? xs1 : xs0);
    }
    wave_hot(tmax, arg);
    const double w_dT = wave_total(s_dT), w_src = wave_total(s_src), w_loss = wave_total(s_loss);
    __shared__ StepStats s_w[kT / kWave];
    if (lane == 0)
        s_w[wave] = StepStats{tmax, arg, ymin, ymax, molten, w_dT, w_src, w_loss};
    __syncthreads();
    if (threadIdx.x == 0) {
        StepStats st = s_w[0];
        PartialAcc b{st.tmax, st.arg, st.ymin, st.ymax};
        for (int i = 1; i < kT / kWave; ++i) {
            b.merge(PartialAcc{s_w[i].tmax, s_w[i].arg, s_w[i].ymin, s_w[i].ymax});
            st.molten += s_w[i].molten;
            st.sum_dT += s_w[i].sum_dT;
            st.sum_src += s_w[i].sum_src;
            st.sum_loss += s_w[i].sum_loss;
        }
        st.tmax = b.tmax; st.arg = b.arg; st.ymin = b.ymin; st.ymax = b.ymax;
        a.part[blockIdx.x] = st;
    }
}

__global__ __launch_bounds__(kT) void k_ext_stats_finalize(const StepStats* __restrict__ part,
                                                           int nwg, int nx, int ny,
                                                           StepStats* __restrict__ out) {
    block_merge_stats(part, nwg, nx, ny, out);
}

// Band height: halved from 32 rows until the grid has about 4k wavefront
// tasks, so small plates still spread over the chip.
int ext_rows(int nx, int ny) {
    const int nwx = (nx + kCols - 1) / kCols;
    int R = kRowsMax;
    while (R > 1 && static_cast<long long>(nwx) * ((ny + R - 1) / R) < 4096) R /= 2;
    return R;
}

int ext_nwg(int nx, int ny) {
    const int R = ext_rows(nx, ny);
    const long long nt = static_cast<long long>((nx + kCols - 1) / kCols) * ((ny + R - 1) / R);
    return static_cast<int>((nt + kT / kWave - 1) / (kT / kWave));
}

bool ext_shape_ok(long long nx, long long ny) {
    return nx >= 2 && ny >= 1 && nx <= (1LL << 28) && ny <= (1LL << 28) &&
           nx * ny < (1LL << 31) - 1;
}

// IN718 table (params.hpp:67-73) and the mushy interval (:63-65).
constexpr double kTp[9] = {298, 473, 673, 873, 1073, 1273, 1473, 1533, 1609};
constexpr double kKp[9] = {8.9, 12.9, 15.8, 18.7, 21.9, 25.6, 28.6, 29.3, 29.6};
constexpr double kCp[9] = {435, 479, 515, 558, 580, 628, 670, 685, 720};
constexpr double kTsol = 1533.0, kTliq = 1609.0, kLF = 210.0e3;
constexpr double kSigma = 5.670374419e-8;   // params.hpp:11

ExtArgs make_ext_args(const float* in, float* out, int ld, const StepConst& k, double T_ref,
                      int bc, double T_bc, double Tm, void* work) {
    ExtArgs a{};
    a.in = in;
    a.out = out;
    a.nx = k.nx;
    a.ny = k.ny;
    a.ld = ld;
    a.R = ext_rows(k.nx, k.ny);
    a.nwx = (k.nx + kCols - 1) / kCols;
    a.ntasks = a.nwx * ((k.ny + a.R - 1) / a.R);
    a.bc = bc;
    a.theta_bc = static_cast<float>(T_bc - T_ref);
    a.theta_m = static_cast<float>(Tm - T_ref);
    a.Tref = static_cast<float>(T_ref);
    a.inv_dx2 = static_cast<float>(k.inv_dx2);
    a.inv_dy2 = static_cast<float>(k.inv_dy2);
    a.inv_h = static_cast<float>(k.inv_h);
    a.dt = static_cast<float>(k.dt);
    a.rho = static_cast<float>(k.rho);
    a.tsub_off = static_cast<float>(T_ref - k.Tsub);
    a.inv_dSub = static_cast<float>(1.0 / k.dSub);
    // q_top(theta) = hConv (T_ref + theta - Tamb) + eps sigma ((T_ref + theta)^4 - Tamb^4)
    const double R0 = T_ref, es = k.radiation ? k.eps * kSigma : 0.0;
    const double Ta4 = k.Tamb * k.Tamb * k.Tamb * k.Tamb;
    a.e0 = static_cast<float>(k.hConv * (R0 - k.Tamb) + es * (R0 * R0 * R0 * R0 - Ta4));
    a.e1 = static_cast<float>(k.hConv + 4.0 * es * R0 * R0 * R0);
    a.e2 = static_cast<float>(6.0 * es * R0 * R0);
    a.e3 = static_cast<float>(4.0 * es * R0);
    a.e4 = static_cast<float>(es);
    a.kmult_m1 = k.kmult > 1.0 ? static_cast<float>(k.kmult - 1.0) : 0.f;
    a.enh_a = static_cast<float>(1.0 / (kTliq - kTsol));
    a.enh_b = static_cast<float>(-kTsol / (kTliq - kTsol));
    a.T_sol = static_cast<float>(kTsol);
    a.T_liq = static_cast<float>(kTliq);
    a.latent = static_cast<float>(kLF / (kTliq - kTsol));
    a.k0 = static_cast<float>(kKp[0]);
    a.cp0 = static_cast<float>(kCp[0]);
    for (int i = 0; i < 8; ++i) {
        const double w = kTp[i + 1] - kTp[i];
        a.rT[i] = static_cast<float>(kTp[i]);
        a.rW[i] = static_cast<float>(w);
        a.kS[i] = static_cast<float>((kKp[i + 1] - kKp[i]) / w);
        a.cS[i] = static_cast<float>((kCp[i + 1] - kCp[i]) / w);
    }
    constexpr double log2e = 1.4426950408889634;
    a.ax = static_cast<float>(log2e * 2.0 * k.dx * k.dx * k.inv_r0_sq);
    a.ay = static_cast<float>(log2e * 2.0 * k.dy * k.dy * k.inv_r0_sq);
    a.cut = static_cast<float>(log2e * 64.0);   // exp(-64): 1.6e-28 of the peak
    a.inv_dx = 1.0 / k.dx;
    a.inv_dy = 1.0 / k.dy;
    a.qS0 = k.peak_S;
    a.part = static_cast<StepStats*>(work);
    return a;
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

extern "C" {

size_t lpbf_f32_ext_workspace_bytes(int nx, int ny) {
    if (!ext_shape_ok(nx, ny)) return 0;
    return static_cast<size_t>(ext_nwg(nx, ny)) * sizeof(StepStats);
}

void lpbf_launch_step_f32_ext(const float* theta_in, float* theta_out, int ld, const StepConst* k,
                              double T_ref, int bc, double T_bc, double Tm,
                              const LaserTick* d_plan, int step, StepStats* d_stats,
                              void* d_work, hipStream_t stream) {
    if (!k || !ext_shape_ok(k->nx, k->ny) || !theta_in || !theta_out || theta_in == theta_out ||
        ld < k->nx || !d_plan || step < 0 || !d_work || k->dt <= 0.0 || k->dSub <= 0.0 ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const ExtArgs a = make_ext_args(theta_in, theta_out, ld, *k, T_ref, bc, T_bc, Tm, d_work);
    const int nwg = ext_nwg(k->nx, k->ny);
    hipLaunchKernelGGL(k_step_f32_ext, dim3(nwg), dim3(kT), 0, stream, a, d_plan, step);
    check_launch();
    if (d_stats) {
        hipLaunchKernelGGL(k_ext_stats_finalize, dim3(1), dim3(kT), 0, stream,
                           static_cast<const StepStats*>(d_work), nwg, k->nx, k->ny, d_stats);
        check_launch();
    }
}

}  // extern "C"
