// SPDX-License-Identifier: 0BSD
// Two-level run with a learned melt-pool window (include/lpbf_multirate.h):
// the coarse far field on the FP32 step kernel, the 128 x 128 fine window
// advanced by the surrogate (residual over the prolonged coarse increment)
// or by the FD step (the window plus a 16-cell margin, in LDS), and the
// transfer kernels between the two grids.  Host driver at the end of the
// file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <vector>

#include "lpbf_common.hpp"
#include "lpbf_f32_kernel.hpp"
#include "lpbf_kernels.h"
#include "lpbf_multirate.h"
#include "lpbf_surrogate.h"

namespace lpbf_amd {
namespace {

constexpr int kW = LPBF_SG_WIN;            // 128 fine cells
constexpr int kWc = kW / 2;                // 64 coarse cells
constexpr int kPix = kW * kW;
constexpr int kCh = LPBF_SG_CH;
constexpr int kT = 256;
constexpr int kSteps = LPBF_SG_STEPS;      // 16 fine steps per block
constexpr int kLead = kW * 3 / 4;          // window cells behind the laser

__device__ inline uint16_t to_bf16(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}

// Cell-centred bilinear prolongation of a coarse plate at fine cell (X, Y):
// the fine centre lies a quarter coarse cell from its coarse centre, towards
// the neighbour (X odd: +1, X even: -1); indices clamp at the plate edge,
// where the mirror edge makes the ghost equal to the edge cell.
__device__ inline float prolong(const float* __restrict__ C, int ldc, int nxc, int nyc, int X,
                                int Y) {
    const int cx = X >> 1, cy = Y >> 1;
    const int nxn = min(max(cx + ((X & 1) ? 1 : -1), 0), nxc - 1);
    const int nyn = min(max(cy + ((Y & 1) ? 1 : -1), 0), nyc - 1);
    const float a = C[static_cast<size_t>(cy) * ldc + cx];
    const float b = C[static_cast<size_t>(cy) * ldc + nxn];
    const float c = C[static_cast<size_t>(nyn) * ldc + cx];
    const float d = C[static_cast<size_t>(nyn) * ldc + nxn];
    return 0.5625f * a + 0.1875f * (b + c) + 0.0625f * d;
}

// prolong() reading the coarse cells from a block of them: cell (u, v) at
// B[(v - by) * ldb + u - bx]; the same weights and clamping at the plate
// edge, so the value equals prolong() of a plate that holds the block.
__device__ inline float prolong_block(const float* __restrict__ B, int ldb, int bx, int by, int nxc,
                                      int nyc, int X, int Y) {
    const int cx = X >> 1, cy = Y >> 1;
    const int nxn = min(max(cx + ((X & 1) ? 1 : -1), 0), nxc - 1);
    const int nyn = min(max(cy + ((Y & 1) ? 1 : -1), 0), nyc - 1);
    const float a = B[(cy - by) * ldb + cx - bx];
    const float b = B[(cy - by) * ldb + nxn - bx];
    const float c = B[(nyn - by) * ldb + cx - bx];
    const float d = B[(nyn - by) * ldb + nxn - bx];
    return 0.5625f * a + 0.1875f * (b + c) + 0.0625f * d;
}

// Conservative quadratic prolongation at fine cell (X, Y), for the field the
// run carries (FD margin, cells the window gains, the field it reports): per
// axis, the quadratic through three neighbouring coarse cell means, averaged
// over the fine half cell, is c + (c_prev - c_next) / 8 on the lower half and
// c + (c_next - c_prev) / 8 on the upper; the 2D value is the tensor product.
// The two halves average to c (a restriction returns the coarse field), and
// a quadratic field is reproduced in cell means.  Indices clamp at the plate
// edge (the mirror ghost).  Behind the laser the trail is a ridge a few
// coarse cells wide, where the bilinear form's error is about ten times
// larger.
__device__ inline float prolong_q(const float* __restrict__ C, int ldc, int nxc, int nyc, int X,
                                  int Y) {
    const int cx = X >> 1, cy = Y >> 1;
    const int xm = max(cx - 1, 0), xp = min(cx + 1, nxc - 1);
    const int ym = max(cy - 1, 0), yp = min(cy + 1, nyc - 1);
    const float sx = (X & 1) ? 0.125f : -0.125f;
    const float sy = (Y & 1) ? 0.125f : -0.125f;
    auto row = [&](int v) {
        const float* r = C + static_cast<size_t>(v) * ldc;
        return fmaf(sx, r[xp] - r[xm], r[cx]);
    };
    const float lo = row(ym), mid = row(cy), hi = row(yp);
    return fmaf(sy, hi - lo, mid);
}

// The 2x2 mean, in one fixed order everywhere (plate, window, residual input).
__device__ inline float mean4(float a, float b, float c, float d) {
    return 0.25f * ((a + b) + (c + d));
}

__global__ __launch_bounds__(kT) void k_mr_restrict(const float* __restrict__ fine, int ldf,
                                                    int nxc, int nyc, float* __restrict__ C,
                                                    int ldc) {
    const long long n = static_cast<long long>(nxc) * nyc;
    for (long long q = static_cast<long long>(blockIdx.x) * kT + threadIdx.x; q < n;
         q += static_cast<long long>(gridDim.x) * kT) {
        const int u = static_cast<int>(q % nxc), v = static_cast<int>(q / nxc);
        const float* r0 = fine + static_cast<size_t>(2 * v) * ldf + 2 * u;
        const float* r1 = r0 + ldf;
        C[static_cast<size_t>(v) * ldc + u] = mean4(r0[0], r0[1], r1[0], r1[1]);
    }
}

// One pixel p of window w's network input; c1v is the coarse field after the
// block, prolonged at the pixel's plate cell.
__device__ inline void gather_pixel(const lpbf_mr_window& W, int w, int p, float c1v,
                                    uint16_t* __restrict__ x0, float* __restrict__ Dout,
                                    float* __restrict__ T0) {
    const int i = p / kW, j = p % kW;
    auto f = [&](int ii, int jj) {
        return W.fine[static_cast<size_t>(W.fy + ii) * W.ldf + W.fx + jj];
    };
    auto R = [&](int ci, int cj) {
        return mean4(f(2 * ci, 2 * cj), f(2 * ci, 2 * cj + 1), f(2 * ci + 1, 2 * cj),
                     f(2 * ci + 1, 2 * cj + 1));
    };
    const float th = f(i, j);
    // fine-scale part: theta minus the prolongation of its own 2x2 means
    // (within the window, indices clamped to its 64 coarse cells)
    const int ci = i >> 1, cj = j >> 1;
    const int ni = min(max(ci + ((i & 1) ? 1 : -1), 0), kWc - 1);
    const int nj = min(max(cj + ((j & 1) ? 1 : -1), 0), kWc - 1);
    const float pr = 0.5625f * R(ci, cj) + 0.1875f * (R(ci, nj) + R(ni, cj)) + 0.0625f * R(ni, nj);
    const float detail = th - pr;
    // coarse increment, prolonged at plate cell (X, Y)
    const int X = W.ox + j, Y = W.oy + i;
    const float dC = c1v - prolong(W.c0, W.ldc, W.nxc, W.nyc, X, Y);
    // laser source map over the block: mean of q * exp(-2 r^2 / r0^2) as
    // exp2 of the cell offsets from the beam in cell units (FP32, far finer
    // than the bf16 channel it feeds)
    const float ax = static_cast<float>(1.4426950408889634 * 2.0 * W.dx * W.dx * W.inv_r0_sq);
    const float ay = static_cast<float>(1.4426950408889634 * 2.0 * W.dy * W.dy * W.inv_r0_sq);
    float src = 0.f;
    for (int s = 0; s < kSteps; ++s) {
        const LaserTick t = W.plan[W.step + s];
        if (t.q_scale == 0.0) continue;
        const float ux = static_cast<float>(X) - static_cast<float>(t.x_l / W.dx);
        const float uy = static_cast<float>(Y) - static_cast<float>(t.y_l / W.dy);
        src += static_cast<float>(t.q_scale) * __builtin_amdgcn_exp2f(-fmaf(ax * ux, ux, ay * uy * uy));
    }
    src *= 1.f / kSteps;
    const size_t q = static_cast<size_t>(w) * kPix + p;
    uint4* out = reinterpret_cast<uint4*>(x0 + q * kCh);
    out[0] = make_uint4(to_bf16(detail * 0.01f) | (uint32_t(to_bf16(th * (1.f / 3000.f))) << 16),
                        to_bf16(dC * 1e-3f) | (uint32_t(to_bf16(src)) << 16),
                        0, 0);
    out[1] = make_uint4(0, 0, 0, 0);
    out[2] = make_uint4(0, 0, 0, 0);
    out[3] = make_uint4(0, 0, 0, 0);
    if (Dout) Dout[q] = dC;
    if (T0) T0[q] = th;
}

__global__ __launch_bounds__(kT) void k_mr_gather(const lpbf_mr_window* __restrict__ win,
                                                  uint16_t* __restrict__ x0,
                                                  float* __restrict__ Dout,
                                                  float* __restrict__ T0) {
    const int w = blockIdx.y;
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= kPix) return;
    const lpbf_mr_window& W = win[w];
    const int X = W.ox + p % kW, Y = W.oy + p / kW;
    const float c1v = W.ld1 > 0 ? prolong_block(W.c1, W.ld1, W.c1x, W.c1y, W.nxc, W.nyc, X, Y)
                                : prolong(W.c1, W.ldc, W.nxc, W.nyc, X, Y);
    gather_pixel(W, w, p, c1v, x0, Dout, T0);
}

__global__ __launch_bounds__(kT) void k_mr_target(const lpbf_mr_window* __restrict__ win,
                                                  const float* __restrict__ T0,
                                                  const float* __restrict__ D,
                                                  float* __restrict__ tgt) {
    const int w = blockIdx.y;
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= kPix) return;
    const lpbf_mr_window& W = win[w];
    const int i = p / kW, j = p % kW;
    const float th1 = W.fine[static_cast<size_t>(W.fy + i) * W.ldf + W.fx + j];
    const size_t q = static_cast<size_t>(w) * kPix + p;
    tgt[q] = ((th1 - T0[q]) - D[q]) * 0.01f;
}

// The FD window: 16 fine steps of the window in one launch.  Workgroup
// (ti, tj) owns the 16 x 16 window cells (16 ti.., 16 tj..) and steps, in LDS,
// the region 16 cells wider on every side (clipped at the plate edge, where
// the clamped neighbour index is the mirror ghost): after 16 steps the
// region's edge error has reached 15 cells in, so the owned cells are exact.
// The region starts from the window's cells and, outside it, the coarse
// field's quadratic prolongation.  Every cell takes the FP32 step kernel's expression
// (theta' = P(theta) + source + cy*(up+down) + cx*(left+right), P the
// centre part kc*theta - loss as one polynomial).
constexpr int kFdTile = 16;
constexpr int kFdReg = kFdTile + 2 * kSteps;   // 48
constexpr int kFdGroups = kT / kFdReg;          // 5 row groups of 48 threads
template <bool RAD>
__global__ __launch_bounds__(kT) void k_mr_fd_window(F32Args a, const float* __restrict__ Wc,
                                                     int ox, int oy,
                                                     const float* __restrict__ C, int ldc,
                                                     int nxc, int nyc,
                                                     const LaserTick* __restrict__ plan,
                                                     int step0, float* __restrict__ out,
                                                     float* __restrict__ md) {
    __shared__ float buf[2][kFdReg * kFdReg];
    if (blockIdx.x == 0 && threadIdx.x == 0 && md) *md = 0.f;
    const int ti = blockIdx.x / (kW / kFdTile), tj = blockIdx.x % (kW / kFdTile);
    const int nx = a.nx, ny = a.ny;
    const int ry0 = max(oy + kFdTile * ti - kSteps, 0);
    const int ry1 = min(oy + kFdTile * ti + kFdTile + kSteps, ny);
    const int rx0 = max(ox + kFdTile * tj - kSteps, 0);
    const int rx1 = min(ox + kFdTile * tj + kFdTile + kSteps, nx);
    const int RH = ry1 - ry0, RW = rx1 - rx0;
    // thread = column c of rows g, g + kFdGroups, ...: no index division
    const int c = threadIdx.x % kFdReg, g = threadIdx.x / kFdReg;
    const bool active = g < kFdGroups && c < RW;
    if (active)
        for (int r = g; r < RH; r += kFdGroups) {
            const int X = rx0 + c, Y = ry0 + r;
            const int wi = Y - oy, wj = X - ox;
            buf[0][r * kFdReg + c] = (wi >= 0 && wi < kW && wj >= 0 && wj < kW)
                                         ? Wc[wi * kW + wj]
                                         : prolong_q(C, ldc, nxc, nyc, X, Y);
        }
    __syncthreads();
    const int cl = max(c - 1, 0), cr = min(c + 1, RW - 1);
    const float ux = static_cast<float>(rx0 + c);
    const float wcut = sqrtf(a.cut / a.ay);
    int cur = 0;
    for (int j = 0; j < kSteps; ++j) {
        const LaserTick t = plan[step0 + j];
        const float xc = static_cast<float>(t.x_l * a.inv_dx);
        const float yc = static_cast<float>(t.y_l * a.inv_dy);
        const float cS = static_cast<float>(a.cS0 * t.q_scale);
        int srlo = static_cast<int>(floorf(yc - wcut)) - 1;
        int srhi = static_cast<int>(ceilf(yc + wcut)) + 1;
        if (t.q_scale == 0.0) { srlo = 1; srhi = 0; }
        const float* A = buf[cur];
        float* B = buf[cur ^ 1];
        if (active)
            for (int r = g; r < RH; r += kFdGroups) {
                const float th = A[r * kFdReg + c];
                const float up = A[max(r - 1, 0) * kFdReg + c];
                const float dn = A[min(r + 1, RH - 1) * kFdReg + c];
                const float lf = A[r * kFdReg + cl];
                const float rt = A[r * kFdReg + cr];
                float P;
                if (RAD) P = fmaf(fmaf(fmaf(fmaf(a.p4, th, a.p3), th, a.p2), th, a.p1), th, a.p0);
                else P = fmaf(a.p1, th, a.p0);
                float v = fmaf(a.cy, up + dn, P);
                const int Y = ry0 + r;
                if (Y >= srlo && Y <= srhi) {
                    const float uy = static_cast<float>(Y) - yc;
                    const float dx = ux - xc;
                    v += cS * __builtin_amdgcn_exp2f(-fmaf(a.ax * dx, dx, a.ay * uy * uy));
                }
                B[r * kFdReg + c] = fmaf(a.cx, lf + rt, v);
            }
        __syncthreads();
        cur ^= 1;
    }
    const float* A = buf[cur];
    for (int q = threadIdx.x; q < kFdTile * kFdTile; q += kT) {
        const int wi = kFdTile * ti + q / kFdTile, wj = kFdTile * tj + q % kFdTile;
        out[wi * kW + wj] = A[(oy + wi - ry0) * kFdReg + (ox + wj - rx0)];
    }
}

// The coarse grid's four steps of a block over the coarse cells the window's
// prolongation reads (kCwOut^2 from (bx, by) = window origin / 2 - 1), for
// the network's input without waiting for the plate's coarse step: the
// region 4 cells wider on every side (clipped at the plate edge) is stepped
// in LDS with the coarse step kernel's expression from the plate at t, with
// the plain coarse coefficients (dt x 4 in one explicit step), which define
// the increment the network was trained to correct.  Cells outside the
// plate are not written (the prolongation never reads them).
constexpr int kCwOut = kWc + 2;                 // 66
constexpr int kCwReg = kCwOut + 2 * 4;          // 74
//
// The step is fused into the network-input gather: workgroup g gathers the
// window's fine rows 2g and 2g + 1 (the window origin is even), whose
// prolongation reads rows g .. g + 2 of the 66 coarse rows above, so it steps
// only those rows and 4 more on either side.  Every such row's four-step
// dependency cone lies in the workgroup's region, and the region is clipped
// at the whole region's edges as above, so the rows hold the same values as
// one workgroup stepping the whole region.
constexpr int kCwRows = 3;                      // coarse rows one workgroup's gather reads
constexpr int kCwRegH = kCwRows + 2 * 4;        // 11
constexpr int kCwGroups = kT / kCwReg;          // 3 row groups of 74 threads
static_assert(kT == 2 * kW, "one fused-gather workgroup per two fine rows");
template <bool RAD>
__global__ __launch_bounds__(kT) void k_mr_gather_cw(F32Args a, const float* __restrict__ C,
                                                     int ldc, int bx, int by,
                                                     const LaserTick* __restrict__ plan,
                                                     int step0,
                                                     const lpbf_mr_window* __restrict__ win,
                                                     uint16_t* __restrict__ x0,
                                                     float* __restrict__ Dout) {
    __shared__ float buf[2][kCwReg * kCwRegH];
    const int nx = a.nx, ny = a.ny;     // coarse plate
    const int g0 = blockIdx.x;          // coarse rows by + g0 .. by + g0 + 2
    const int ry0 = max(by + g0 - 4, 0), ry1 = min(by + g0 + kCwRows + 4, ny);
    const int rx0 = max(bx - 4, 0), rx1 = min(bx + kCwOut + 4, nx);
    const int RH = ry1 - ry0, RW = rx1 - rx0;
    // thread = column c of rows g, g + kCwGroups, ...
    const int c = threadIdx.x % kCwReg, g = threadIdx.x / kCwReg;
    const bool active = g < kCwGroups && c < RW;
    // every load of the region is issued before the first one is waited for
    // (a loop storing each row to LDS as it arrives would pay the memory
    // latency once per row)
    constexpr int kCwRowsPer = (kCwRegH + kCwGroups - 1) / kCwGroups;   // 4
    float ld[kCwRowsPer];
#pragma unroll
    for (int q = 0; q < kCwRowsPer; ++q) {
        const int r = g + q * kCwGroups;
        ld[q] = (active && r < RH) ? C[static_cast<size_t>(ry0 + r) * ldc + rx0 + c] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < kCwRowsPer; ++q) {
        const int r = g + q * kCwGroups;
        if (active && r < RH) buf[0][r * kCwReg + c] = ld[q];
    }
    __syncthreads();
    const float wcut = sqrtf(a.cut / a.ay);
    const int cl = max(c - 1, 0), cr = min(c + 1, RW - 1);
    const float ux = static_cast<float>(rx0 + c);
    int cur = 0;
    for (int j = 0; j < 4; ++j) {
        const LaserTick t = plan[step0 + j];
        const float xc = static_cast<float>(t.x_l * a.inv_dx);
        const float yc = static_cast<float>(t.y_l * a.inv_dy);
        const float cS = static_cast<float>(a.cS0 * t.q_scale);
        int srlo = static_cast<int>(floorf(yc - wcut)) - 1;
        int srhi = static_cast<int>(ceilf(yc + wcut)) + 1;
        if (t.q_scale == 0.0) { srlo = 1; srhi = 0; }
        const float* A = buf[cur];
        float* B = buf[cur ^ 1];
        if (active)
            for (int r = g; r < RH; r += kCwGroups) {
                const float th = A[r * kCwReg + c];
                const float up = A[max(r - 1, 0) * kCwReg + c];
                const float dn = A[min(r + 1, RH - 1) * kCwReg + c];
                const float lf = A[r * kCwReg + cl];
                const float rt = A[r * kCwReg + cr];
                float P;
                if (RAD) P = fmaf(fmaf(fmaf(fmaf(a.p4, th, a.p3), th, a.p2), th, a.p1), th, a.p0);
                else P = fmaf(a.p1, th, a.p0);
                float v = fmaf(a.cy, up + dn, P);
                const int Y = ry0 + r;
                if (Y >= srlo && Y <= srhi) {
                    const float uy = static_cast<float>(Y) - yc;
                    const float dx = ux - xc;
                    v += cS * __builtin_amdgcn_exp2f(-fmaf(a.ax * dx, dx, a.ay * uy * uy));
                }
                B[r * kCwReg + c] = fmaf(a.cx, lf + rt, v);
            }
        __syncthreads();
        cur ^= 1;
    }
    // the gather of fine rows 2 g0 and 2 g0 + 1, prolonging from the region
    const float* A = buf[cur];
    const lpbf_mr_window& W = win[0];
    const int p = g0 * kT + threadIdx.x;
    const int X = W.ox + p % kW, Y = W.oy + p / kW;
    const float c1v = prolong_block(A, kCwReg, rx0, ry0, W.nxc, W.nyc, X, Y);
    gather_pixel(W, 0, p, c1v, x0, Dout, nullptr);
}

// Window after the block from the network: base (window + prolonged coarse
// increment) plus the residual, weighted by the seam ramp; with the FD
// window F, the largest |network - FD| over the cells at least `overlap`
// from the edge.
__global__ __launch_bounds__(kT) void k_mr_compose(const float* __restrict__ Wc,
                                                   const float* __restrict__ D,
                                                   const float* __restrict__ y, int overlap,
                                                   float* __restrict__ Wt,
                                                   const float* __restrict__ F,
                                                   float* __restrict__ md) {
    const int p = blockIdx.x * kT + threadIdx.x;
    const int i = p / kW, j = p % kW;
    const int e = min(min(i, j), min(kW - 1 - i, kW - 1 - j));
    const float a = fminf(1.f, static_cast<float>(e) / static_cast<float>(overlap));
    const float v = (Wc[p] + D[p]) + a * (100.f * y[p]);
    Wt[p] = v;
    if (F) {
        float d = e >= overlap ? fabsf(v - F[p]) : 0.f;
        d = wave_max_f(d);
        if ((threadIdx.x & 63) == 0) atomicMax(reinterpret_cast<int*>(md), __float_as_int(d));
    }
}

// After a block: the window's new cells (src) overwrite their coarse cells
// (2x2 means, all but m coarse cells at each window edge), and the window
// moves to its next origin: cells it keeps come from src, the others are
// prolonged (quadratic) from the coarse grid.  The two parts are independent:
// a cell outside the old window is prolonged from coarse cells at most one cell
// inside the old window's coarse region, which lie in the unwritten edge
// band (m >= 1).  src is either a materialised window (Ws) or, with Ws
// null, the network result Wc + D + ramp * 100 y, formed on the fly.
struct PostSrc {
    const float* Ws;
    const float* Wc;
    const float* D;
    const float* y;
    int overlap;
    __device__ float operator()(int i, int j) const {
        const int p = i * kW + j;
        if (Ws) return Ws[p];
        const int e = min(min(i, j), min(kW - 1 - i, kW - 1 - j));
        const float a = fminf(1.f, static_cast<float>(e) / static_cast<float>(overlap));
        return (Wc[p] + D[p]) + a * (100.f * y[p]);
    }
};

__global__ __launch_bounds__(kT) void k_mr_post(PostSrc src, int ox0, int oy0, int ox1, int oy1,
                                                int m, float* __restrict__ C, int ldc, int nxc,
                                                int nyc, float* __restrict__ Wn) {
    const int q = blockIdx.x * kT + threadIdx.x;
    if (q < kPix) {
        const int i = q / kW, j = q % kW;
        const int X = ox1 + j, Y = oy1 + i;
        const int wi = Y - oy0, wj = X - ox0;
        Wn[q] = (wi >= 0 && wi < kW && wj >= 0 && wj < kW) ? src(wi, wj)
                                                          : prolong_q(C, ldc, nxc, nyc, X, Y);
        return;
    }
    const int n = kWc - 2 * m;
    const int r = q - kPix;
    if (r >= n * n) return;
    const int v = m + r / n, u = m + r % n;
    C[static_cast<size_t>(oy0 / 2 + v) * ldc + ox0 / 2 + u] =
        mean4(src(2 * v, 2 * u), src(2 * v, 2 * u + 1), src(2 * v + 1, 2 * u),
              src(2 * v + 1, 2 * u + 1));
}

__global__ __launch_bounds__(kT) void k_mr_field(const float* __restrict__ Wc, int ox, int oy,
                                                 const float* __restrict__ C, int ldc, int nxc,
                                                 int nyc, double T_ref, double* __restrict__ T,
                                                 int nx, int ny) {
    const long long n = static_cast<long long>(nx) * ny;
    for (long long q = static_cast<long long>(blockIdx.x) * kT + threadIdx.x; q < n;
         q += static_cast<long long>(gridDim.x) * kT) {
        const int X = static_cast<int>(q % nx), Y = static_cast<int>(q / nx);
        const int wi = Y - oy, wj = X - ox;
        const float v = (wi >= 0 && wi < kW && wj >= 0 && wj < kW)
                            ? Wc[wi * kW + wj]
                            : prolong_q(C, ldc, nxc, nyc, X, Y);
        T[q] = T_ref + static_cast<double>(v);
    }
}

// g(f(theta)) for quartics g, f, terms above theta^4 dropped.
void compose4(const double g[5], const double f[5], double out[5]) {
    double pw[5] = {1.0, 0.0, 0.0, 0.0, 0.0};   // f^d
    for (int i = 0; i < 5; ++i) out[i] = 0.0;
    for (int d = 0; d <= 4; ++d) {
        for (int i = 0; i < 5; ++i) out[i] += g[d] * pw[i];
        double nx[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; i + j < 5; ++j) nx[i + j] += pw[i] * f[j];
        for (int i = 0; i < 5; ++i) pw[i] = nx[i];
    }
}

// The coarse step's coefficients, chosen so that one coarse step (dt x 4,
// dx x 2) reproduces four fine explicit steps rather than one explicit step
// of four times the length.  With the fine step theta' = Phi(theta) + D theta
// (Phi(theta) = theta - dt loss(theta) pointwise, D the diffusion), four fine
// steps are Phi^4(theta) + 4 Phi'^3 D theta + O(D^2): the coarse centre part
// is Phi^4 (composed in FP64; the dropped terms above theta^4 come to
// 5e-6 K at 3000 K) and its diffusion numbers carry the factor
// (1 - dt loss'(0))^3.  A plain coarse step's loss term, 1 - 4 lambda dt
// against (1 - lambda dt)^4, drifts by 6 (lambda dt)^2 per coarse step: about
// 4 K over the millisecond a hot track takes to cool on a 50 um plate.
F32Coeffs coarse_coeffs(const StepConst& kf, double T_ref) {
    double L[5];
    loss_poly(kf, T_ref, L);
    const double phi[5] = {-L[0], 1.0 - L[1], -L[2], -L[3], -L[4]};
    double phi2[5], phi4[5];
    compose4(phi, phi, phi2);
    compose4(phi2, phi2, phi4);
    const double g = (1.0 - L[1]) * (1.0 - L[1]) * (1.0 - L[1]);
    F32Coeffs c;
    // the coarse grid's diffusion numbers equal the fine grid's (dt x 4, dx^2 x 4)
    c.cx = kf.alpha * kf.dt * kf.inv_dx2 * g;
    c.cy = kf.alpha * kf.dt * kf.inv_dy2 * g;
    for (int i = 0; i < 5; ++i) c.p[i] = phi4[i];
    c.p[1] -= 2.0 * (c.cx + c.cy);
    return c;
}

int grid_for(long long n) {
    return static_cast<int>(std::min<long long>((n + kT - 1) / kT, 8192));
}

struct Failure {
    int code;
};

void hip_ok(hipError_t e) {
    if (e != hipSuccess) throw Failure{static_cast<int>(e)};
}

void lib_ok() {
    const int e = lpbf_get_last_error();
    if (e) throw Failure{e};
}

}  // namespace
}  // namespace lpbf_amd

using namespace lpbf_amd;

struct lpbf_mr_run {
    lpbf_mr_config cfg{};
    StepConst kf{}, kc{};
    int nx = 0, ny = 0, nxc = 0, nyc = 0, ldc = 0, n_blocks = 0, depth_c = 4;
    std::vector<int> ox, oy;             // window origin per block (n_blocks + 1)
    F32Args af{};                        // the fine step's constants
    F32Args ac{};                        // the plain coarse step's constants (network input)
    F32Coeffs ccoef{};                   // the plate's coarse step (four fine steps)
    hipStream_t s = nullptr;             // the caller's stream: the window
    hipStream_t s2 = nullptr;            // the coarse plate's steps
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    const void* pack = nullptr;
    float* C[2] = {nullptr, nullptr};
    float* C2 = nullptr;                 // third coarse buffer (coarse depth < 4)
    float* Wb[2] = {nullptr, nullptr};
    float* Wt = nullptr;
    float* Wf = nullptr;                 // the FD window after a block
    float* D = nullptr;
    float* y = nullptr;
    void* x0 = nullptr;
    void* sg_work = nullptr;
    void* work_c = nullptr;
    float* md = nullptr;
    float* md_host = nullptr;            // pinned
    LaserTick* d_plan = nullptr;
    LaserTick* d_cplan = nullptr;
    lpbf_mr_window* d_win = nullptr;     // per block, both coarse parities fixed by the block
    int cc = 0, wc = 0, block = 0;
    int fallback_left = 0, since_check = 0;
    lpbf_mr_counters cnt{};
    std::vector<float> diffs;

    ~lpbf_mr_run() {
        for (void* p : {static_cast<void*>(C[0]), static_cast<void*>(C[1]), static_cast<void*>(C2),
                        static_cast<void*>(Wb[0]), static_cast<void*>(Wb[1]),
                        static_cast<void*>(Wt), static_cast<void*>(Wf), static_cast<void*>(D),
                        static_cast<void*>(y), x0, sg_work, work_c, static_cast<void*>(md),
                        static_cast<void*>(d_plan), static_cast<void*>(d_cplan),
                        static_cast<void*>(d_win)})
            if (p) (void)hipFree(p);
        if (md_host) (void)hipHostFree(md_host);
        if (ev_a) (void)hipEventDestroy(ev_a);
        if (ev_b) (void)hipEventDestroy(ev_b);
        if (s2) (void)hipStreamDestroy(s2);
    }

    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        hip_ok(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)));
        return static_cast<T*>(p);
    }

    void fd_window(int b) {
        if (af.a4 != 0.f)
            hipLaunchKernelGGL(k_mr_fd_window<true>, dim3((kW / kFdTile) * (kW / kFdTile)),
                               dim3(kT), 0, s, af, Wb[wc], ox[b], oy[b], C[cc], ldc, nxc, nyc,
                               d_plan, kSteps * b, Wf, md);
        else
            hipLaunchKernelGGL(k_mr_fd_window<false>, dim3((kW / kFdTile) * (kW / kFdTile)),
                               dim3(kT), 0, s, af, Wb[wc], ox[b], oy[b], C[cc], ldc, nxc, nyc,
                               d_plan, kSteps * b, Wf, md);
    }

    // Four coarse steps from C[cc] into C[cc ^ 1], leaving C[cc] (time t)
    // intact for the increment: one launch at depth 4, otherwise an even
    // number of launches alternating through the third buffer.
    void coarse_step(int b) {
        std::vector<int> ds;   // depth 4: {4}; 3: {3, 1}; 2: {2, 2}; 1: {1, 1, 1, 1}
        for (int t = 0; t < 4; t += depth_c) ds.push_back(std::min(depth_c, 4 - t));
        const float* in = C[cc];
        int t = 0;
        for (size_t l = 0; l < ds.size(); ++l) {
            float* out = (l + 1 == ds.size() || l % 2) ? C[cc ^ 1] : C2;
            // the statistics switch is off: Tm only feeds the molten count
            step_f32_launch_coeffs(in, out, ldc, kc, cfg.T_ref, ccoef, d_cplan, 4 * b + t, ds[l],
                                   cfg.coarse_rows, work_c, s2);
            in = out;
            t += ds[l];
        }
    }

    // The network input of block b (and its coarse increment D), with the
    // window-local coarse step computed inside the gather.
    void gather_window(int b) {
        const int bx = ox[b] / 2 - 1, by = oy[b] / 2 - 1;
        if (ac.a4 != 0.f)
            hipLaunchKernelGGL(k_mr_gather_cw<true>, dim3(kPix / kT), dim3(kT), 0, s, ac, C[cc],
                               ldc, bx, by, d_cplan, 4 * b, d_win + b,
                               static_cast<uint16_t*>(x0), D);
        else
            hipLaunchKernelGGL(k_mr_gather_cw<false>, dim3(kPix / kT), dim3(kT), 0, s, ac, C[cc],
                               ldc, bx, by, d_cplan, 4 * b, d_win + b,
                               static_cast<uint16_t*>(x0), D);
        check_launch();
    }

    // One block.  The coarse plate's four steps run on s2 while the window
    // (FD window, or window-local coarse step + network) runs on s; the post
    // kernel, which writes the window's restriction into the new coarse
    // field, waits for both.
    void step_block() {
        const int b = block;
        const bool net = cfg.use_network && fallback_left == 0;
        const bool recheck = net && since_check == 0;
        const bool fd = !net || recheck;
        const float* Wc = Wb[wc];
        float* Wn = Wb[wc ^ 1];
        hip_ok(hipEventRecord(ev_a, s));          // C[cc] holds time t
        hip_ok(hipStreamWaitEvent(s2, ev_a, 0));
        coarse_step(b);
        hip_ok(hipEventRecord(ev_b, s2));
        if (fd) fd_window(b);
        const float* src = nullptr;
        bool keep_fd = !net;
        if (net) {
            gather_window(b);
            lpbf_sg_forward(pack, x0, 1, sg_work, y, s);
            ++cnt.calls;
            src = nullptr;   // formed on the fly by the post kernel
            if (recheck) {
                hipLaunchKernelGGL(k_mr_compose, dim3(kPix / kT), dim3(kT), 0, s, Wc, D, y,
                                   cfg.overlap, Wt, Wf, md);
                src = Wt;
                hip_ok(hipMemcpyAsync(md_host, md, sizeof(float), hipMemcpyDeviceToHost, s));
                hip_ok(hipStreamSynchronize(s));
                lib_ok();
                const float d = *md_host;
                diffs.push_back(d);
                ++cnt.rechecks;
                cnt.max_recheck_K = std::max(cnt.max_recheck_K, static_cast<double>(d));
                if (!(d <= cfg.tol_K)) {
                    ++cnt.fallbacks;
                    keep_fd = true;
                    fallback_left = cfg.recheck_every - 1;
                    since_check = 0;
                } else {
                    cnt.max_applied_recheck_K =
                        std::max(cnt.max_applied_recheck_K, static_cast<double>(d));
                }
            }
            if (!keep_fd) {
                ++cnt.applied;
                since_check = (since_check + 1) % cfg.recheck_every;
            }
        } else if (fallback_left > 0) {
            --fallback_left;
        }
        if (keep_fd) {
            src = Wf;
            ++cnt.fd_blocks;
        }
        hip_ok(hipStreamWaitEvent(s, ev_b, 0));   // the coarse plate at t + 16
        const int m = cfg.overlap / 2;
        const int nr = kWc - 2 * m;
        PostSrc ps{src, Wc, D, y, cfg.overlap};
        hipLaunchKernelGGL(k_mr_post, dim3((kPix + nr * nr + kT - 1) / kT), dim3(kT), 0, s, ps,
                           ox[b], oy[b], ox[b + 1], oy[b + 1], m, C[cc ^ 1], ldc, nxc, nyc, Wn);
        check_launch();
        cc ^= 1;
        wc ^= 1;
        ++block;
        ++cnt.blocks;
    }
};

namespace {

int fail(int code) {
    record_error(code);
    return code;
}

}  // namespace

extern "C" {

void lpbf_mr_restrict(const float* fine, int ldf, int nx, int ny, float* coarse, int ldc,
                      hipStream_t stream) {
    if (!fine || !coarse || nx < 2 || ny < 2 || (nx & 1) || (ny & 1) || ldf < nx ||
        ldc < nx / 2) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const long long n = static_cast<long long>(nx / 2) * (ny / 2);
    hipLaunchKernelGGL(k_mr_restrict, dim3(grid_for(n)), dim3(kT), 0, stream, fine, ldf, nx / 2,
                       ny / 2, coarse, ldc);
    check_launch();
}

void lpbf_mr_gather(const lpbf_mr_window* d_win, int n, void* d_x0, float* d_D, float* d_T0,
                    hipStream_t stream) {
    if (!d_win || n < 1 || n > 65535 || !d_x0) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_mr_gather, dim3(kPix / kT, n), dim3(kT), 0, stream, d_win,
                       static_cast<uint16_t*>(d_x0), d_D, d_T0);
    check_launch();
}

void lpbf_mr_target(const lpbf_mr_window* d_win, int n, const float* d_T0, const float* d_D,
                    float* d_target, hipStream_t stream) {
    if (!d_win || n < 1 || n > 65535 || !d_T0 || !d_D || !d_target) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    hipLaunchKernelGGL(k_mr_target, dim3(kPix / kT, n), dim3(kT), 0, stream, d_win, d_T0, d_D,
                       d_target);
    check_launch();
}

int lpbf_mr_window_origins(const LaserTick* h_plan, int n_blocks, int nx, int ny, double dx,
                           double dy, int* ox, int* oy) {
    if (!h_plan || n_blocks < 0 || nx < kW || ny < kW || !(dx > 0.0) || !(dy > 0.0) || !ox ||
        !oy)
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    int dir = 1;   // +x until the plan shows motion in x
    for (int b = 0; b <= n_blocks; ++b) {
        const LaserTick& t = h_plan[kSteps * b];
        if (b < n_blocks) {
            const double d = h_plan[kSteps * (b + 1)].x_l - t.x_l;
            if (d > 0.0) dir = 1;
            else if (d < 0.0) dir = -1;
        }
        // the laser at 3/4 of the window along its travel, centred across it
        const int lx = static_cast<int>(std::floor(t.x_l / dx + 0.5));
        const int ly = static_cast<int>(std::floor(t.y_l / dy + 0.5));
        int x = lx - (dir > 0 ? kLead : kW - kLead);
        int y = ly - kW / 2;
        ox[b] = std::min(std::max(x, 0), nx - kW) & ~1;
        oy[b] = std::min(std::max(y, 0), ny - kW) & ~1;
    }
    return 0;
}

int lpbf_mr_coarse_coeffs(const StepConst* k, double T_ref, double* out7) {
    if (!k || !out7 || !(k->dt > 0.0) || !(k->dx > 0.0) || !(k->dy > 0.0))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    const F32Coeffs c = coarse_coeffs(*k, T_ref);
    out7[0] = c.cx;
    out7[1] = c.cy;
    for (int i = 0; i < 5; ++i) out7[2 + i] = c.p[i];
    return 0;
}

int lpbf_mr_create(const lpbf_mr_config* cfg, const LaserTick* h_plan, const void* d_pack,
                   hipStream_t stream, lpbf_mr_run** out) {
    if (!cfg || !h_plan || !out) return fail(LPBF_ERR_INVALID_ARGUMENT);
    *out = nullptr;
    const StepConst& k = cfg->k;
    if (k.nx < 2 * kW || k.ny < 2 * kW || (k.nx & 1) || (k.ny & 1) || k.nx > (1 << 16) ||
        k.ny > (1 << 16) || cfg->n_steps < kSteps || cfg->n_steps % kSteps ||
        cfg->recheck_every < 1 || cfg->overlap < 2 || cfg->overlap > 32 || (cfg->overlap & 1) ||
        !(cfg->tol_K >= 0.0) || (cfg->use_network && !d_pack) || cfg->coarse_depth < 0 ||
        cfg->coarse_depth > 4 || cfg->coarse_rows < 0 || !(k.dt > 0.0) || !(k.dx > 0.0) ||
        !(k.dy > 0.0))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    lpbf_mr_run* r = new (std::nothrow) lpbf_mr_run;
    if (!r) return fail(static_cast<int>(hipErrorOutOfMemory));
    try {
        r->cfg = *cfg;
        r->s = stream;
        r->pack = d_pack;
        r->kf = k;
        r->nx = k.nx;
        r->ny = k.ny;
        r->nxc = k.nx / 2;
        r->nyc = k.ny / 2;
        r->ldc = (r->nxc + 3) & ~3;
        r->depth_c = cfg->coarse_depth;
        // coarse grid: 2x2 cells, four steps; alpha dt / dx^2 is unchanged
        StepConst kc = k;
        kc.nx = r->nxc;
        kc.ny = r->nyc;
        kc.dx = 2.0 * k.dx;
        kc.dy = 2.0 * k.dy;
        kc.inv_dx2 = 1.0 / (kc.dx * kc.dx);
        kc.inv_dy2 = 1.0 / (kc.dy * kc.dy);
        kc.dt = 4.0 * k.dt;
        r->kc = kc;
        const int nb = cfg->n_steps / kSteps;
        r->n_blocks = nb;
        // window origins
        r->ox.resize(static_cast<size_t>(nb) + 1);
        r->oy.resize(static_cast<size_t>(nb) + 1);
        if (lpbf_mr_window_origins(h_plan, nb, k.nx, k.ny, k.dx, k.dy, r->ox.data(),
                                   r->oy.data()))
            throw Failure{LPBF_ERR_INVALID_ARGUMENT};
        r->af = make_args(nullptr, nullptr, k.nx, k, cfg->T_ref, LPBF_BC_MIRROR, cfg->T_ref,
                          cfg->T_ref + 1e9, 1, 0, nullptr);
        // the window-local coarse step (the network's coarse increment) keeps
        // the plain coarse step the network was trained on; the plate's
        // coarse step takes the four-fine-step coefficients
        r->ac = make_args(nullptr, nullptr, r->ldc, kc, cfg->T_ref, LPBF_BC_MIRROR, cfg->T_ref,
                          cfg->T_ref + 1e9, 1, 0, nullptr);
        r->ccoef = coarse_coeffs(k, cfg->T_ref);
        hip_ok(hipStreamCreateWithFlags(&r->s2, hipStreamNonBlocking));
        hip_ok(hipEventCreateWithFlags(&r->ev_a, hipEventDisableTiming));
        hip_ok(hipEventCreateWithFlags(&r->ev_b, hipEventDisableTiming));
        // coarse ticks: one per four fine steps, the beam's mean strength at
        // its strength-weighted mean position, shifted because coarse cell u
        // is centred on fine position 2u + 1/2 while the kernel places it at 2u
        std::vector<LaserTick> cplan(4 * static_cast<size_t>(nb));
        for (int q = 0; q < 4 * nb; ++q) {
            double qs = 0.0, xs = 0.0, ys = 0.0, xm = 0.0, ym = 0.0;
            for (int j = 0; j < 4; ++j) {
                const LaserTick& t = h_plan[4 * q + j];
                qs += t.q_scale;
                xs += t.q_scale * t.x_l;
                ys += t.q_scale * t.y_l;
                xm += 0.25 * t.x_l;
                ym += 0.25 * t.y_l;
            }
            LaserTick c;
            c.x_l = (qs > 0.0 ? xs / qs : xm) - 0.5 * k.dx;
            c.y_l = (qs > 0.0 ? ys / qs : ym) - 0.5 * k.dy;
            c.q_scale = 0.25 * qs;
            cplan[q] = c;
        }
        const size_t np = static_cast<size_t>(cfg->n_steps) + kSteps;
        r->d_plan = r->alloc<LaserTick>(np);
        r->d_cplan = r->alloc<LaserTick>(cplan.size());
        hip_ok(hipMemcpyAsync(r->d_plan, h_plan, sizeof(LaserTick) * np, hipMemcpyHostToDevice,
                              stream));
        hip_ok(hipMemcpyAsync(r->d_cplan, cplan.data(), sizeof(LaserTick) * cplan.size(),
                              hipMemcpyHostToDevice, stream));
        const size_t cbytes = static_cast<size_t>(r->ldc) * r->nyc;
        for (int q = 0; q < 2; ++q) {
            r->C[q] = r->alloc<float>(cbytes);
            hip_ok(hipMemsetAsync(r->C[q], 0, sizeof(float) * cbytes, stream));
        }
        if (r->depth_c == 0) {
            // one launch of four coarse steps per block: time the band
            // heights at depth 4 (the autotuner's per-step pick at other
            // depths costs a second launch per block)
            static const int kRows[] = {0, 8, 12, 16, 20, 24, 32, 48};
            size_t wb = 0;
            for (int R : kRows)
                if (R == 0 || R <= r->nyc) wb = std::max(wb, lpbf_f32_workspace_bytes(r->nxc, r->nyc, 4, R));
            void* work = r->alloc<uint8_t>(wb);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            double best = 1e300;
            int best_rows = 0;
            try {
                hip_ok(hipEventCreate(&e0));
                hip_ok(hipEventCreate(&e1));
                for (int R : kRows) {
                    if (R > r->nyc) continue;
                    auto go = [&] {
                        step_f32_launch_coeffs(r->C[0], r->C[1], r->ldc, r->kc, cfg->T_ref,
                                               r->ccoef, r->d_cplan, 0, 4, R, work, stream);
                    };
                    go();
                    hip_ok(hipEventRecord(e0, stream));
                    for (int rep = 0; rep < 5; ++rep) go();
                    hip_ok(hipEventRecord(e1, stream));
                    hip_ok(hipEventSynchronize(e1));
                    float ms = 0.f;
                    hip_ok(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) { best = ms; best_rows = R; }
                }
                lib_ok();
            } catch (...) {
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                (void)hipFree(work);
                throw;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            (void)hipFree(work);
            hip_ok(hipMemsetAsync(r->C[1], 0, sizeof(float) * cbytes, stream));
            r->depth_c = 4;
            r->cfg.coarse_rows = best_rows;
        }
        for (int q = 0; q < 2; ++q) {
            r->Wb[q] = r->alloc<float>(kPix);
            hip_ok(hipMemsetAsync(r->Wb[q], 0, sizeof(float) * kPix, stream));
        }
        if (r->depth_c < 4) r->C2 = r->alloc<float>(cbytes);
        r->Wt = r->alloc<float>(kPix);
        r->Wf = r->alloc<float>(kPix);
        r->D = r->alloc<float>(kPix);
        r->y = r->alloc<float>(kPix);
        r->x0 = r->alloc<uint16_t>(static_cast<size_t>(kPix) * kCh);
        if (cfg->use_network) {
            const size_t wb = lpbf_sg_work_bytes(1, 0);
            if (!wb) throw Failure{LPBF_ERR_INVALID_ARGUMENT};
            r->sg_work = r->alloc<uint8_t>(wb);
        }
        r->work_c = r->alloc<uint8_t>(
            lpbf_f32_workspace_bytes(r->nxc, r->nyc, r->depth_c, r->cfg.coarse_rows));
        r->md = r->alloc<float>(1);
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&r->md_host), sizeof(float),
                             hipHostMallocDefault));
        // network windows: block b reads window buffer b % 2 and coarse
        // buffers b % 2 (time t) and (b + 1) % 2 (t + 16)
        std::vector<lpbf_mr_window> win(static_cast<size_t>(nb));
        for (int b = 0; b < nb; ++b) {
            lpbf_mr_window& w = win[b];
            w = lpbf_mr_window{};
            w.fine = r->Wb[b & 1];
            w.c0 = r->C[b & 1];
            // c1 stays empty: the run's gather steps the coarse cells it reads
            w.plan = r->d_plan;
            w.dx = k.dx;
            w.dy = k.dy;
            w.inv_r0_sq = k.inv_r0_sq;
            w.ldf = kW;
            w.fx = w.fy = 0;
            w.ox = r->ox[b];
            w.oy = r->oy[b];
            w.ldc = r->ldc;
            w.nxc = r->nxc;
            w.nyc = r->nyc;
            w.step = kSteps * b;
        }
        r->d_win = r->alloc<lpbf_mr_window>(win.size());
        hip_ok(hipMemcpyAsync(r->d_win, win.data(), sizeof(lpbf_mr_window) * win.size(),
                              hipMemcpyHostToDevice, stream));
        hip_ok(hipStreamSynchronize(stream));
        lib_ok();
    } catch (Failure const& e) {
        delete r;
        return fail(e.code);
    } catch (std::bad_alloc const&) {
        delete r;
        return fail(static_cast<int>(hipErrorOutOfMemory));
    }
    *out = r;
    return 0;
}

void lpbf_mr_destroy(lpbf_mr_run* r) {
    if (r) {
        (void)hipStreamSynchronize(r->s);
        if (r->s2) (void)hipStreamSynchronize(r->s2);
    }
    delete r;
}

int lpbf_mr_advance(lpbf_mr_run* r, int n_blocks) {
    if (!r || n_blocks < 0) return fail(LPBF_ERR_INVALID_ARGUMENT);
    try {
        for (int q = 0; q < n_blocks && r->block < r->n_blocks; ++q) r->step_block();
        lib_ok();
    } catch (Failure const& e) {
        return fail(e.code);
    }
    return 0;
}

int lpbf_mr_field(const lpbf_mr_run* r, double* d_T) {
    if (!r || !d_T) return fail(LPBF_ERR_INVALID_ARGUMENT);
    const long long n = static_cast<long long>(r->nx) * r->ny;
    hipLaunchKernelGGL(k_mr_field, dim3(grid_for(n)), dim3(kT), 0, r->s, r->Wb[r->wc],
                       r->ox[r->block], r->oy[r->block], r->C[r->cc], r->ldc, r->nxc, r->nyc,
                       r->cfg.T_ref, d_T, r->nx, r->ny);
    check_launch();
    return lpbf_get_last_error();
}

int lpbf_mr_get_counters(const lpbf_mr_run* r, lpbf_mr_counters* out) {
    if (!r || !out) return fail(LPBF_ERR_INVALID_ARGUMENT);
    *out = r->cnt;
    return 0;
}

int lpbf_mr_recheck_diffs(const lpbf_mr_run* r, float* out, int max) {
    if (!r || max < 0 || (max > 0 && !out)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return -LPBF_ERR_INVALID_ARGUMENT;
    }
    const int n = std::min<int>(max, static_cast<int>(r->diffs.size()));
    std::copy(r->diffs.begin(), r->diffs.begin() + n, out);
    return static_cast<int>(r->diffs.size());
}

}  // extern "C"
