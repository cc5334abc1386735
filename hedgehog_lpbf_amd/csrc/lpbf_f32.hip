// SPDX-License-Identifier: 0BSD
// FP32 performance path for gfx950: temporally blocked explicit-Euler steps
// with fused per-step statistics.
//
// Field: theta = T - T_ref in FP32, rows of pitch `ld` floats (ld % 4 == 0,
// 16-byte aligned base).  Storing the excess temperature keeps small
// increments near ambient above FP32 resolution.
//
// Schedule (MI355X-first):
//   * a wavefront owns a strip of 256 loaded columns (lane l holds 4
//     contiguous columns as one 16-byte load) that overlaps its neighbours by
//     kOv = 4 columns on each side, so 248 columns are written per strip;
//   * the strip streams rows top to bottom through a K-level register
//     pipeline: loading row i advances level j to row i-j, so one pass over
//     HBM performs K time steps (K <= 4 keeps every written column valid,
//     since the invalid overlap grows by one column per level);
//   * x neighbours are cross-lane shuffles, y neighbours the register window;
//     edge conditions are fixed up per row, only in edge strips;
//   * per level, every owned cell-update feeds the step's statistics (peak T
//     with its index, molten rows and count, energy sums): FP32 per lane,
//     FP64 across lanes and workgroups;
//   * four strips form a 256-thread workgroup; workgroup ids are remapped so
//     consecutive tiles land on the same XCD (shared halo cache lines in L2).
#include <hip/hip_runtime.h>

#include "lpbf_common.hpp"
#include "lpbf_kernels.h"

namespace lpbf_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kOv = 4;                          // overlap columns per side
constexpr int kLoadCols = 4 * kWave;            // 256
constexpr int kOutCols = kLoadCols - 2 * kOv;   // 248
constexpr int kMaxDepth = 4;
constexpr int kXcds = 8;
constexpr int kPf = 2;                          // row prefetch distance

struct F32Args {
    const float* in;
    float* out;
    int nx, ny, ld;
    int nstrips, R, ntasks, nwg;
    int bc;
    float cx, cy;        // alpha*dt/dx^2, alpha*dt/dy^2
    float cHB, c0;       // linear loss per step: cHB*theta + c0
    float cR;            // radiation loss per step / (T^4 - Ta^4); 0 if off
    float dA;            // T_ref - Tamb
    float cTT;           // T_ref + Tamb
    float Tref;
    float Ta2;           // Tamb^2
    float theta_m;       // Tm - T_ref
    float theta_bc;      // T_bc - T_ref
    float ax, ay;        // log2(e) * 2*dx^2/r0^2, log2(e) * 2*dy^2/r0^2
    float cut;           // exponent (base 2) beyond which the source is dropped
    double inv_dx, inv_dy, cS0;  // 1/dx, 1/dy, dt*peak_S/(rho_cp*h)
    double Tref_d;
    StepStats* part;     // [depth][nwg]
};

__device__ inline int xcd_remap(int bid, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int xcd = bid % kXcds, local = bid / kXcds;
    return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

struct LevelAcc {
    float bv;       // best theta
    int bi;         // its row-major index
    int ymin, ymax;
    int molten;
    float sd, ss, sl;
};

template <int K, int BC>
__global__ __launch_bounds__(kThreads) void k_step_f32(F32Args a,
                                                       const LaserTick* __restrict__ plan,
                                                       int step0) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int wg = xcd_remap(blockIdx.x, a.nwg);
    const int task = wg * kWaves + wave;
    const int nx = a.nx, ny = a.ny;

    float xc[K], yc[K], cS[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const LaserTick t = plan[step0 + j];
        xc[j] = static_cast<float>(t.x_l * a.inv_dx);
        yc[j] = static_cast<float>(t.y_l * a.inv_dy);
        cS[j] = static_cast<float>(a.cS0 * t.q_scale);
    }

    LevelAcc acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = LevelAcc{-3.0e38f, 0x7fffffff, ny, -1, 0, 0.f, 0.f, 0.f};

    if (task < a.ntasks) {
        const int band = task / a.nstrips, strip = task % a.nstrips;
        const int L0 = strip * kOutCols - kOv;
        const int col4 = L0 + 4 * lane;
        const bool ld_ok = col4 >= 0 && col4 < nx;
        const bool own_lane = lane >= 1 && lane <= kWave - 2 && col4 < nx;
        bool own_e[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) own_e[e] = own_lane && (col4 + e < nx);
        const int ya = band * a.R;
        const int yb = min(ya + a.R, ny);
        const int y_lo = max(ya - K, 0);
        const int y_hi = min(yb + K, ny);
        const int i_end = yb - 1 + K;   // level-K row yb-1 is produced at i = yb-1+K
        // edge columns held by this strip
        const bool has_left = (L0 < 0);                  // column -1 at lane 0, elem 3
        const int pR = nx - L0;                          // position of column nx
        const bool has_right = (pR >= 0 && pR < kLoadCols);
        const int padl = pR >> 2, pade = pR & 3;
        const int pL = nx - 1 - L0;                      // position of column nx-1
        const bool has_last = (pL >= 0 && pL < kLoadCols);
        const int lastl = pL >> 2, laste = pL & 3;

        // Mirror: column -1 := column 0, column nx := column nx-1.
        // Dirichlet: column 0 and nx-1 := theta_bc (levels >= 1).
        auto fix_mirror = [&](float4& v) {
            if (has_left) {
                const float t = __shfl_down(v.x, 1);
                if (lane == 0) v.w = t;
            }
            if (has_right) {
                const float t = __shfl_up(v.w, 1);
                if (lane == padl) {
                    if (pade == 0) v.x = t;
                    else if (pade == 1) v.y = v.x;
                    else if (pade == 2) v.z = v.y;
                    else v.w = v.z;
                }
            }
        };
        auto fix_dirichlet = [&](float4& v) {
            if (has_left && lane == 1) v.x = a.theta_bc;   // column 0 at position 4
            if (has_last && lane == lastl) {
                if (laste == 0) v.x = a.theta_bc;
                else if (laste == 1) v.y = a.theta_bc;
                else if (laste == 2) v.z = a.theta_bc;
                else v.w = a.theta_bc;
            }
        };

        float4 W[K][3];
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
            for (int q = 0; q < 3; ++q) W[j][q] = make_float4(0.f, 0.f, 0.f, 0.f);

        // rows are prefetched kPf iterations ahead of their use
        auto load_row = [&](int i) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < y_hi && ld_ok)
                v = *reinterpret_cast<const float4*>(a.in + static_cast<long long>(i) * a.ld + col4);
            return v;
        };
        float4 pf[kPf];
#pragma unroll
        for (int q = 0; q < kPf; ++q) pf[q] = load_row(y_lo + q);

        for (int i = y_lo; i <= i_end; ++i) {
            float4 v = pf[0];
#pragma unroll
            for (int q = 0; q < kPf - 1; ++q) pf[q] = pf[q + 1];
            pf[kPf - 1] = load_row(i + kPf);
            if (BC == LPBF_BC_MIRROR && i < y_hi) fix_mirror(v);
            W[0][0] = W[0][1]; W[0][1] = W[0][2]; W[0][2] = v;

#pragma unroll
            for (int j = 1; j <= K; ++j) {
                const int r = i - j;
                const int lo = max(ya - (K - j), 0);
                const int hi = min(yb + (K - j), ny);
                float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
                if (r >= lo && r < hi) {
                    const float4 c = W[j - 1][1];
                    const float4 up = (r == 0) ? c : W[j - 1][0];
                    const float4 dn = (r == ny - 1) ? c : W[j - 1][2];
                    const float wl = __shfl_up(c.w, 1);    // column left of c.x
                    const float wr = __shfl_down(c.x, 1);  // column right of c.w
                    const float cv[4] = {c.x, c.y, c.z, c.w};
                    const float xm[4] = {wl, c.x, c.y, c.z};
                    const float xp[4] = {c.y, c.z, c.w, wr};
                    const float uv[4] = {up.x, up.y, up.z, up.w};
                    const float dv[4] = {dn.x, dn.y, dn.z, dn.w};
                    // source: the whole row is skipped unless it is near the laser
                    const float uy = static_cast<float>(r) - yc[j - 1];
                    const float ey = a.ay * uy * uy;
                    const bool src_row = ey < a.cut;
                    float out4[4], srcv[4], lossv[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float th = cv[e];
                        float s = 0.f;
                        if (src_row) {
                            const float ux = static_cast<float>(col4 + e) - xc[j - 1];
                            s = cS[j - 1] * exp2f(-(a.ax * ux * ux + ey));
                        }
                        const float Tt = th + a.Tref;
                        float loss = a.cHB * th + a.c0;
                        if (a.cR != 0.f) {
                            const float rad = (th + a.dA) * (th + a.cTT) * (Tt * Tt + a.Ta2);
                            loss += a.cR * rad;
                        }
                        const float lap = a.cx * ((xm[e] + xp[e]) - 2.f * th)
                                        + a.cy * ((uv[e] + dv[e]) - 2.f * th);
                        out4[e] = th + ((lap + s) - loss);
                        srcv[e] = s;
                        lossv[e] = loss;
                    }
                    res = make_float4(out4[0], out4[1], out4[2], out4[3]);
                    if (BC == LPBF_BC_DIRICHLET) {
                        if (r == 0 || r == ny - 1)
                            res = make_float4(a.theta_bc, a.theta_bc, a.theta_bc, a.theta_bc);
                        else
                            fix_dirichlet(res);
                    } else {
                        fix_mirror(res);
                    }
                    if (r >= ya && r < yb) {
                        if (j == K && own_lane)
                            *reinterpret_cast<float4*>(a.out + static_cast<long long>(r) * a.ld + col4) = res;
                        const float rv[4] = {res.x, res.y, res.z, res.w};
                        LevelAcc& A = acc[j - 1];
                        const bool upd_row = !(BC == LPBF_BC_DIRICHLET && (r == 0 || r == ny - 1));
                        int m = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            if (!own_e[e]) continue;
                            if (rv[e] > A.bv) { A.bv = rv[e]; A.bi = r * nx + col4 + e; }
                            m += (rv[e] > a.theta_m) ? 1 : 0;
                            const bool upd = upd_row &&
                                !(BC == LPBF_BC_DIRICHLET && (col4 + e == 0 || col4 + e == nx - 1));
                            if (upd) {
                                A.sd += rv[e] - cv[e];
                                A.ss += srcv[e];
                                A.sl += lossv[e];
                            }
                        }
                        if (m) {
                            A.molten += m;
                            if (r < A.ymin) A.ymin = r;
                            A.ymax = r;
                        }
                    }
                }
                if (j < K) { W[j][0] = W[j][1]; W[j][1] = W[j][2]; W[j][2] = res; }
            }
        }
    }

    // per level: wavefront then workgroup merge -> a.part[j*nwg + wg]
    __shared__ PartialAcc s_pa[kWaves];
    __shared__ double s_sum[3][kWaves];
    __shared__ long long s_m[kWaves];
    const long long N = static_cast<long long>(nx) * ny;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const LevelAcc& A = acc[j];
        PartialAcc p;
        p.tmax = (A.bi == 0x7fffffff) ? -1.0 : a.Tref_d + static_cast<double>(A.bv);
        p.arg = (A.bi == 0x7fffffff) ? N : static_cast<long long>(A.bi);
        p.ymin = A.ymin;
        p.ymax = A.ymax;
        const PartialAcc w = wave_reduce(p);
        const double wd = wave_sum(static_cast<double>(A.sd));
        const double ws = wave_sum(static_cast<double>(A.ss));
        const double wl = wave_sum(static_cast<double>(A.sl));
        const long long wm = wave_sum(static_cast<long long>(A.molten));
        if (lane == 0) {
            s_pa[wave] = w; s_sum[0][wave] = wd; s_sum[1][wave] = ws; s_sum[2][wave] = wl;
            s_m[wave] = wm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            PartialAcc b = s_pa[0];
            StepStats st;
            st.molten = s_m[0]; st.sum_dT = s_sum[0][0]; st.sum_src = s_sum[1][0];
            st.sum_loss = s_sum[2][0];
            for (int q = 1; q < kWaves; ++q) {
                b.merge(s_pa[q]);
                st.molten += s_m[q]; st.sum_dT += s_sum[0][q];
                st.sum_src += s_sum[1][q]; st.sum_loss += s_sum[2][q];
            }
            st.tmax = b.tmax; st.arg = b.arg; st.ymin = b.ymin; st.ymax = b.ymax;
            a.part[static_cast<long long>(j) * a.nwg + wg] = st;
        }
        __syncthreads();
    }
}

// One workgroup per level merges that level's workgroup records.
__global__ __launch_bounds__(kThreads) void k_stats_finalize_levels(const StepStats* __restrict__ part,
                                                                    int nwg, int nx, int ny,
                                                                    StepStats* __restrict__ out) {
    block_merge_stats(part + static_cast<long long>(blockIdx.x) * nwg, nwg, nx, ny,
                      out + blockIdx.x);
}

__global__ void k_noise_f32(float* __restrict__ th, int nx, int ny, int ld,
                            unsigned long long seed, float base, float amp2) {
    const long long n = static_cast<long long>(nx) * ny;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(i / nx), x = static_cast<int>(i % nx);
        th[static_cast<long long>(y) * ld + x] =
            base + amp2 * static_cast<float>(noise_u(seed, static_cast<unsigned long long>(i)));
    }
}

__global__ void k_theta_to_T(const float* __restrict__ th, int ld, int nx, int ny, double Tref,
                             double* __restrict__ T) {
    const long long n = static_cast<long long>(nx) * ny;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(i / nx), x = static_cast<int>(i % nx);
        T[i] = Tref + static_cast<double>(th[static_cast<long long>(y) * ld + x]);
    }
}

__global__ void k_T_to_theta(const double* __restrict__ T, int nx, int ny, double Tref,
                             float* __restrict__ th, int ld) {
    const long long n = static_cast<long long>(nx) * ny;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(i / nx), x = static_cast<int>(i % nx);
        th[static_cast<long long>(y) * ld + x] = static_cast<float>(T[i] - Tref);
    }
}

int rows_per_band(int nx, int ny, int depth) {
    // enough wavefront tiles to fill 256 CUs several times, long enough
    // bands that the 2*depth recomputed halo rows stay a small fraction
    const int nstrips = (nx + kOutCols - 1) / kOutCols;
    int R = (depth > 1) ? 128 : 64;
    while (R > 16 && static_cast<long long>(nstrips) * ((ny + R - 1) / R) < 4096) R /= 2;
    if (depth > 1 && R < 8 * depth) R = 8 * depth;
    return R;
}

struct Geometry {
    int nstrips, R, ntasks, nwg;
};

Geometry geometry(int nx, int ny, int depth, int rows) {
    Geometry g;
    g.nstrips = (nx + kOutCols - 1) / kOutCols;
    g.R = rows > 0 ? rows : rows_per_band(nx, ny, depth);
    g.ntasks = g.nstrips * ((ny + g.R - 1) / g.R);
    g.nwg = (g.ntasks + kWaves - 1) / kWaves;
    return g;
}

template <int K>
void launch_k(const F32Args& a, const LaserTick* plan, int step0, hipStream_t s) {
    if (a.bc == LPBF_BC_DIRICHLET)
        hipLaunchKernelGGL((k_step_f32<K, LPBF_BC_DIRICHLET>), dim3(a.nwg), dim3(kThreads), 0, s, a,
                           plan, step0);
    else
        hipLaunchKernelGGL((k_step_f32<K, LPBF_BC_MIRROR>), dim3(a.nwg), dim3(kThreads), 0, s, a,
                           plan, step0);
}

// The FP32 kernels index cells with 32-bit ints and rely on nx, ny for every
// address; shapes outside this range are rejected.
bool f32_shape_ok(long long nx, long long ny) {
    return nx >= 2 && ny >= 1 && nx <= (1LL << 28) && ny <= (1LL << 28) &&
           nx * ny < (1LL << 31) - 1;
}

F32Args make_args(const float* in, float* out, int ld, const StepConst& k, double T_ref, int bc,
                  double T_bc, double Tm, int depth, int rows, void* work) {
    F32Args a{};
    const Geometry g = geometry(k.nx, k.ny, depth, rows);
    a.in = in; a.out = out; a.nx = k.nx; a.ny = k.ny; a.ld = ld;
    a.nstrips = g.nstrips; a.R = g.R; a.ntasks = g.ntasks; a.nwg = g.nwg;
    a.bc = bc;
    const double dt = k.dt;
    a.cx = static_cast<float>(k.alpha * dt * k.inv_dx2);
    a.cy = static_cast<float>(k.alpha * dt * k.inv_dy2);
    // loss*dt = dt*((hConv*(T-Tamb) + (K/dSub)*(T-Tsub))*inv_h/rho_cp + rad)
    const double cH = dt * k.hConv * k.inv_h / k.rho_cp;
    const double cB = dt * (k.K / k.dSub) * k.inv_h / k.rho_cp;
    a.cHB = static_cast<float>(cH + cB);
    a.c0 = static_cast<float>(cH * (T_ref - k.Tamb) + cB * (T_ref - k.Tsub));
    a.cR = k.radiation ? static_cast<float>(dt * k.eps * 5.670374419e-8 * k.inv_h / k.rho_cp) : 0.f;
    a.dA = static_cast<float>(T_ref - k.Tamb);
    a.cTT = static_cast<float>(T_ref + k.Tamb);
    a.Tref = static_cast<float>(T_ref);
    a.Ta2 = static_cast<float>(k.Tamb * k.Tamb);
    a.theta_m = static_cast<float>(Tm - T_ref);
    a.theta_bc = static_cast<float>(T_bc - T_ref);
    const double log2e = 1.4426950408889634;
    a.ax = static_cast<float>(log2e * 2.0 * k.dx * k.dx * k.inv_r0_sq);
    a.ay = static_cast<float>(log2e * 2.0 * k.dy * k.dy * k.inv_r0_sq);
    a.cut = static_cast<float>(log2e * 64.0);   // exp(-64) ~ 1.6e-28 of the peak
    a.inv_dx = 1.0 / k.dx;
    a.inv_dy = 1.0 / k.dy;
    a.cS0 = dt * k.peak_S / (k.rho_cp * k.h);
    a.Tref_d = T_ref;
    a.part = static_cast<StepStats*>(work);
    return a;
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

extern "C" {

size_t lpbf_f32_workspace_bytes(int nx, int ny, int depth, int rows) {
    if (!f32_shape_ok(nx, ny) || depth < 1 || depth > kMaxDepth || rows < 0) return 0;
    // enough for every depth up to `depth` (a run ends with shorter launches)
    size_t need = 0;
    for (int d = 1; d <= depth; ++d) {
        const Geometry g = geometry(nx, ny, d, rows);
        const size_t b = sizeof(StepStats) * static_cast<size_t>(g.nwg) * static_cast<size_t>(d);
        if (b > need) need = b;
    }
    return need;
}

void lpbf_launch_step_f32(const float* theta_in, float* theta_out, int ld, const StepConst* k,
                          double T_ref, int bc, double T_bc, double Tm, const LaserTick* d_plan,
                          int step0, int depth, int rows, StepStats* d_stats, void* d_work,
                          hipStream_t stream) {
    if (!k || !f32_shape_ok(k->nx, k->ny) || !theta_in || !theta_out || theta_in == theta_out ||
        ld < k->nx || (ld & 3) || (reinterpret_cast<uintptr_t>(theta_in) & 15) ||
        (reinterpret_cast<uintptr_t>(theta_out) & 15) || !d_plan || step0 < 0 || depth < 1 ||
        depth > kMaxDepth || rows < 0 || !d_work ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const F32Args a = make_args(theta_in, theta_out, ld, *k, T_ref, bc, T_bc, Tm, depth, rows,
                                d_work);
    switch (depth) {
        case 1: launch_k<1>(a, d_plan, step0, stream); break;
        case 2: launch_k<2>(a, d_plan, step0, stream); break;
        case 3: launch_k<3>(a, d_plan, step0, stream); break;
        default: launch_k<4>(a, d_plan, step0, stream); break;
    }
    check_launch();
    if (d_stats) {
        hipLaunchKernelGGL(k_stats_finalize_levels, dim3(depth), dim3(kThreads), 0, stream, a.part,
                           a.nwg, k->nx, k->ny, d_stats);
        check_launch();
    }
}

void lpbf_launch_f32_stats_finalize(const void* d_work, int nx, int ny, int depth, int rows,
                                    StepStats* d_stats, hipStream_t stream) {
    if (!d_work || !d_stats || !f32_shape_ok(nx, ny) || depth < 1 || depth > kMaxDepth ||
        rows < 0) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Geometry g = geometry(nx, ny, depth, rows);
    hipLaunchKernelGGL(k_stats_finalize_levels, dim3(depth), dim3(kThreads), 0, stream,
                       static_cast<const StepStats*>(d_work), g.nwg, nx, ny, d_stats);
    check_launch();
}

void lpbf_launch_noise_f32(float* theta, int nx, int ny, int ld, unsigned long long seed,
                           double base, double amp, hipStream_t stream) {
    if (!theta || !f32_shape_ok(nx, ny) || ld < nx) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const long long n = static_cast<long long>(nx) * ny;
    const long long want = (n + kThreads - 1) / kThreads;
    const int grid = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_noise_f32, dim3(grid), dim3(kThreads), 0, stream, theta, nx, ny, ld, seed,
                       static_cast<float>(base), static_cast<float>(2.0 * amp));
    check_launch();
}

void lpbf_launch_theta_to_T(const float* theta, int ld, int nx, int ny, double T_ref, double* T,
                            hipStream_t stream) {
    if (!theta || !T || !f32_shape_ok(nx, ny) || ld < nx) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const long long n = static_cast<long long>(nx) * ny;
    const long long want = (n + kThreads - 1) / kThreads;
    const int grid = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_theta_to_T, dim3(grid), dim3(kThreads), 0, stream, theta, ld, nx, ny, T_ref,
                       T);
    check_launch();
}

void lpbf_launch_T_to_theta(const double* T, int nx, int ny, double T_ref, float* theta, int ld,
                            hipStream_t stream) {
    if (!theta || !T || !f32_shape_ok(nx, ny) || ld < nx) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const long long n = static_cast<long long>(nx) * ny;
    const long long want = (n + kThreads - 1) / kThreads;
    const int grid = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_T_to_theta, dim3(grid), dim3(kThreads), 0, stream, T, nx, ny, T_ref, theta,
                       ld);
    check_launch();
}

}  // extern "C"
