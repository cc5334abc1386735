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
//   * x neighbours are whole-wavefront DPP shifts folded into the adds, y
//     neighbours the register window; mirror edges on widths that are a
//     multiple of 4 use a reflected halo (no per-level edge work), other
//     edge conditions are fixed up per level, only in edge strips;
//   * per level, every owned cell-update feeds the step's statistics (peak T
//     with its index, molten rows and count, energy sums): peak and sums per
//     lane (FP32, the theta sum FP64; the peak's row staged in LDS), molten
//     rows and count per wavefront from ballots, FP64 across lanes and
//     workgroups;
//   * four strips form a 256-thread workgroup; workgroup ids are remapped so
//     consecutive tiles land on the same XCD (shared halo cache lines in L2).
//
// This header holds the tile body shared by the single-field kernel
// (lpbf_f32.hip) and the multi-instance kernel (lpbf_f32_batch.hip), and the
// host-side tiling geometry and argument set-up both use.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

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
constexpr int kPf = 2;                          // row prefetch distance

struct F32Args {
    const float* in;
    float* out;
    int nx, ny, ld;
    int nstrips, R, ntasks, nwg;
    int bc;
    float cx, cy;          // alpha*dt/dx^2, alpha*dt/dy^2
    float kc;              // 1 - 2*cx - 2*cy, the centre weight of the update
    // dt*loss(theta) in K = (((a4*th + a3)*th + a2)*th + a1)*th + a0 (Horner):
    // convection + substrate are linear in theta, radiation expands
    // (T_ref + theta)^4 - Tamb^4 around T_ref.
    float a0, a1, a2, a3, a4;
    // the centre part of the update, kc*theta - dt*loss(theta), as one
    // polynomial: p0 = -a0, p1 = kc - a1, p2 = -a2, p3 = -a3, p4 = -a4
    float p0, p1, p2, p3, p4;
    float theta_m;         // Tm - T_ref
    float theta_bc;        // T_bc - T_ref
    float ax, ay;          // log2(e) * 2*dx^2/r0^2, log2(e) * 2*dy^2/r0^2
    float cut;             // base-2 exponent beyond which the source is dropped
    double inv_dx, inv_dy, cS0;  // 1/dx, 1/dy, dt*peak_S/(rho_cp*h)
    double Tref_d;
    double kc_d;           // kc in FP64: the loss sum is kc * sum(theta) - sum(centre part)
    StepStats* part;       // [depth][nwg]
};

// Largest of four finite values: v_max3 + v_max, without the NaN-quieting
// v_max x,x the compiler inserts before fmaxf of values it cannot prove
// canonical (no operand here can be a NaN).
__device__ __forceinline__ float max4(float4 v) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3\n\tv_max_f32 %0, %0, %4"
        : "=&v"(r) : "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
    return r;
}

// Whole-wavefront lane shifts (DPP wave_shr:1 / wave_shl:1, GFX9 family):
// lane l receives lane l-1 (from_left) or lane l+1 (from_right); the lane
// without a source receives 0.
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// Empty register constraint: marks the value as produced at this point
// (emits no instruction).
template <typename T>
__device__ __forceinline__ void pin(T& v) {
    asm volatile("" : "+v"(v));
}

__device__ __forceinline__ float& elem(float4& v, int e) {
    return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}

// v[e] = t for a run-time e, as four selects (a reference into the vector
// with a run-time index would put it in scratch memory).
__device__ __forceinline__ void set_elem(float4& v, int e, float t) {
    v.x = e == 0 ? t : v.x;
    v.y = e == 1 ? t : v.y;
    v.z = e == 2 ? t : v.z;
    v.w = e == 3 ? t : v.w;
}

// Row loads and stores through device (global address space) pointers: the
// multi-instance kernel reads its buffer pointers from memory, where the
// compiler cannot infer the address space and would emit flat accesses,
// which also count on the scalar-load counter, so that every wait for an
// instance argument would wait for the row prefetches as well.
// (a native vector type: float4's copy constructor would bind a generic
// reference and undo the address space)
typedef float f32x4 __attribute__((ext_vector_type(4)));
using gf32x4 = __attribute__((address_space(1))) f32x4;
__device__ __forceinline__ float4 ld_row4(const float* p) {
    const f32x4 t = *(const gf32x4*)p;
    return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void st_row4(float* p, float4 v) {
    const f32x4 t = {v.x, v.y, v.z, v.w};
    *(gf32x4*)p = t;
}

struct LevelAcc {
    float bv;        // best theta of this lane
    int bi;          // the body row it was found in; after the row loop its
                     // row-major index (smallest on ties)
    int ymin, ymax;  // molten rows of the wavefront's owned cells (wave-uniform)
    int molten;      // molten owned cells of the wavefront (wave-uniform)
    double S;        // sum of theta over the owned cells after this step
    float SP;        // sum of the step's centre parts kc*theta - dt*loss
    float SS;        // sum of dt*S
};

// Workgroup `wg` of one field: four strip tiles, K steps, and the
// workgroup's per-level statistics records a.part[j*a.nwg + wg].  The field
// moves from a.in to a.out, or, with BUFS, from src to dst: the
// multi-instance kernel passes its buffers directly and keeps `a` in memory,
// where its fields are loaded at their uses (a local copy of `a` costs nine
// more VGPRs and one wave per SIMD at K = 4; the kernel-argument form is
// allocated best reading a.in / a.out itself).
// RAGGED = false promises nx % 4 == 0: strip origins are multiples of 4, so
// no lane then holds both column nx-1 and a column >= nx, and the masking of
// such columns out of the statistics drops from the code.
template <int K, int BC, bool RAD, bool BUFS = false, bool RAGGED = true>
__device__ __forceinline__ void step_tile(const F32Args& a, const float* __restrict__ src,
                                          float* __restrict__ dst,
                                          const LaserTick* __restrict__ plan, int step0, int wg) {
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform: everything derived from the tile index lives in SGPRs
    // and the row/strip conditions below are scalar branches
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int task = wg * kWaves + wave;
    const int nx = a.nx, ny = a.ny;
    constexpr float kNegInf = -3.0e38f;
    // Mirror edges use a reflected halo: the ghost columns -4..-1 and
    // nx..nx+3 enter as the mirror images of columns 3..0 and nx-1..nx-4,
    // and the ghost cells are updated like any other (the source at the
    // mirrored column).  The update is symmetric, so every level's ghost
    // column -1 equals column 0 bit for bit (and nx equals nx-1): the mirror
    // condition holds without per-level fix-ups.  Each level spoils one more
    // ghost column from the far side, so four stay exact for K <= 4.  The
    // left ghosts are one lane (its load reversed); on the right, a width
    // that is a multiple of 4 puts them in one lane too, and otherwise they
    // straddle two lanes and are filled once per loaded row from the lanes
    // holding columns nx-1..nx-4 (fill_right).
    constexpr bool kReflect = (BC == LPBF_BC_MIRROR);

    // per level: laser position in cell units, source scale, and the rows
    // [srlo, srhi] whose source can matter (|r - yc| below the cut-off, plus
    // one row of margin); the row test is then a scalar compare.  A strip
    // whose columns all lie beyond the cut-off in x gets no source rows.
    // The Gaussian is separable: per level each lane holds its columns'
    // x factors gx = 2^(-ax ux^2) (and their sum over its owned columns), so
    // a source row costs one exponential for the row's y factor.
    float xc[K], yc[K], cS[K];
    int srlo[K], srhi[K];
    const float wcut = sqrtf(a.cut / a.ay);
    const float wcut_x = sqrtf(a.cut / a.ax);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const LaserTick t = plan[step0 + j];
        // uniform values: keep them in SGPRs
        xc[j] = __int_as_float(__builtin_amdgcn_readfirstlane(
            __float_as_int(static_cast<float>(t.x_l * a.inv_dx))));
        yc[j] = __int_as_float(__builtin_amdgcn_readfirstlane(
            __float_as_int(static_cast<float>(t.y_l * a.inv_dy))));
        cS[j] = __int_as_float(__builtin_amdgcn_readfirstlane(
            __float_as_int(static_cast<float>(a.cS0 * t.q_scale))));
        srlo[j] = __builtin_amdgcn_readfirstlane(static_cast<int>(floorf(yc[j] - wcut)) - 1);
        srhi[j] = __builtin_amdgcn_readfirstlane(static_cast<int>(ceilf(yc[j] + wcut)) + 1);
        if (t.q_scale == 0.0) { srlo[j] = 1; srhi[j] = 0; }   // beam off: no source rows
    }

    // per level and lane: the four values of the row holding the lane's peak
    __shared__ float4 s_best[K][kWaves][kWave];
    LevelAcc acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = LevelAcc{kNegInf, 0x7fffffff, ny, -1, 0, 0.0, 0.f, 0.f};
    double S0 = 0.0;   // sum of the input theta over the lane's cells
    // Lanes 0 and 63 (strip halo) and lanes past column nx own no cells.
    // Every lane accumulates its own peak and sums; the ones that own nothing
    // are dropped once, before the wavefront reductions (the molten counts
    // are masked with own_mask as they are taken).
    bool owner = false;
    unsigned long long own_mask = 0;   // the owning lanes (wave-uniform)

    if (task < a.ntasks) {
        const int band = task / a.nstrips, strip = task % a.nstrips;
        const int L0 = strip * kOutCols - kOv;
        const int col4 = L0 + 4 * lane;
        float gx[K][4], gxs[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            // strip columns [L0, L0 + kLoadCols) against |x - xc| <= wcut_x
            // (one column of margin)
            if (static_cast<float>(L0 + kLoadCols) < xc[j] - wcut_x - 1.f ||
                static_cast<float>(L0) > xc[j] + wcut_x + 1.f) {
                srlo[j] = 1;
                srhi[j] = 0;
            }
            gxs[j] = 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // with the reflected halo, ghost columns take the source of
                // the column they mirror
                int cm = col4 + e;
                if (kReflect) {
                    // twice: on a width below 4 a mirrored column can itself
                    // be a ghost
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        cm = cm < 0 ? -1 - cm : (cm >= nx ? 2 * nx - 1 - cm : cm);
                }
                const float ux = static_cast<float>(cm) - xc[j];
                // v_exp_f32: results below 2^-126 flush to 0, far under
                // FP32 resolution of any theta
                gx[j][e] = __builtin_amdgcn_exp2f(-(a.ax * ux) * ux);
                // the source sum covers the owned columns (lanes past column
                // nx own nothing and are dropped later; only a straddling
                // lane needs the per-column test)
                const int cc = col4 + e;
                if (BC == LPBF_BC_DIRICHLET) gxs[j] += (cc >= 1 && cc < nx - 1) ? gx[j][e] : 0.f;
                else gxs[j] += (!RAGGED || cc < nx) ? gx[j][e] : 0.f;
            }
        }
        const bool ld_ok = col4 >= 0 && col4 < nx;
        const bool own_lane = lane >= 1 && lane <= kWave - 2 && col4 < nx;
        owner = own_lane;
        own_mask = __builtin_amdgcn_ballot_w64(own_lane);
        // the polynomial's leading coefficient in a VGPR: its first FMA
        // would otherwise read two scalar operands, one more than a VALU
        // instruction may, and take a copy on every row and level
        float lead;
        asm volatile("v_mov_b32 %0, %1" : "=v"(lead) : "s"(RAD ? a.p4 : a.p1));
        const int ya = band * a.R;
        const int yb = min(ya + a.R, ny);
        const int y_lo = max(ya - K, 0);
        const int y_hi = min(yb + K, ny);
        const int i_end = yb - 1 + K;   // level-K row yb-1 is produced at i = yb-1+K
        const bool has_left = (L0 < 0);                  // column -1 at lane 0, elem 3
        const int pR = nx - L0;                          // position of column nx
        const bool has_right = (pR >= 0 && pR < kLoadCols);
        const int padl = pR >> 2, pade = pR & 3;
        const int pL = nx - 1 - L0;                      // position of column nx-1
        const bool has_last = (pL >= 0 && pL < kLoadCols);
        const int lastl = pL >> 2, laste = pL & 3;
        // the lane that holds column nx-1 may also hold columns >= nx
        const bool ragged = RAGGED && has_right && pade != 0;
        // strips holding a domain edge column take the column fix-ups
        const bool edge = has_left || has_right || has_last;

        // Mirror: column -1 := column 0, column nx := column nx-1.
        auto fix_mirror = [&](float4& v) {
            if (has_left) {
                const float t = from_right(v.x);
                if (lane == 0) v.w = t;
            }
            if (has_right) {
                const float t = from_left(v.w);
                if (lane == padl) {
                    if (pade == 0) v.x = t;
                    else if (pade == 1) v.y = v.x;
                    else if (pade == 2) v.z = v.y;
                    else v.w = v.z;
                }
            }
        };
        // Reflected halo on a ragged right edge: column nx sits at element
        // pade (1..3) of lane padl, so the ghosts nx..nx+3 := nx-1..nx-4 lie
        // in lanes padl and padl+1 and their sources in lanes padl-1 and
        // padl (one or two lanes to the left).
        auto fill_right = [&](float4& v) {
            const bool L1 = lane == padl, L2 = lane == padl + 1;
            if (pade == 1) {
                // padl: x=nx-1 y=nx z=nx+1 w=nx+2; padl+1: x=nx+3
                const float bw = from_left(v.w), bz = from_left(v.z);
                const float cy = from_left(from_left(v.y));
                if (L1) { v.y = v.x; v.z = bw; v.w = bz; }
                if (L2) v.x = cy;
            } else if (pade == 2) {
                // padl: z=nx w=nx+1; padl+1: x=nx+2 y=nx+3
                const float cw = from_left(from_left(v.w)), cz = from_left(from_left(v.z));
                if (L1) { v.z = v.y; v.w = v.x; }
                if (L2) { v.x = cw; v.y = cz; }
            } else {
                // padl: w=nx; padl+1: x=nx+1 y=nx+2 z=nx+3
                const float by = from_left(v.y), bx = from_left(v.x);
                const float cw = from_left(from_left(v.w));
                if (L1) v.w = v.z;
                if (L2) { v.x = by; v.y = bx; v.z = cw; }
            }
        };
        // Dirichlet: columns 0 and nx-1 := theta_bc (dir_cols keeps them out
        // of the sums).
        auto fix_dirichlet = [&](float4& v) {
            if (has_left && lane == 1) v.x = a.theta_bc;   // column 0
            if (has_last && lane == lastl) set_elem(v, laste, a.theta_bc);
        };
        // Dirichlet edge cells are held, not updated, and stay out of the
        // energy sums: columns 0 and nx-1 drop out of a row's theta sum
        auto dir_cols = [&](float4 v) {
            if (BC == LPBF_BC_DIRICHLET && edge) {
                if (has_left && lane == 1) v.x = 0.f;
                if (has_last && lane == lastl) set_elem(v, laste, 0.f);
            }
            return v;
        };
        // A row's contribution to the lane's sums; on the lane holding column
        // nx-1, columns >= nx drop out.
        auto row_sum = [&](float4 v) {
            if (ragged && lane == lastl) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (e > laste) elem(v, e) = 0.f;
            }
            return (v.x + v.y) + (v.z + v.w);
        };

        // reflected halo: lanes left of column 0 or right of column nx-1 load
        // the nearest in-field quad (reversed when the row is consumed)
        // nx4 - 4: the last quad that starts at or before column nx-1 (nx4 is
        // nx rounded up to a multiple of 4, at most the pitch)
        const int nx4 = (nx + 3) & ~3;
        const int col4c = kReflect ? min(max(col4, 0), nx4 - 4) : min(max(col4, 0), a.ld - 4);
        const bool rev = kReflect && (col4 < 0 || col4 >= nx);
        auto load_row = [&](int i) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (kReflect) {
                if (i < y_hi)
                    v = ld_row4((BUFS ? src : a.in) + static_cast<long long>(i) * a.ld + col4c);
            } else if (i < y_hi && ld_ok) {
                v = ld_row4((BUFS ? src : a.in) + static_cast<long long>(i) * a.ld + col4);
            }
            return v;
        };
        // rows past y_hi are never consumed, so the row index is clamped
        // instead of tested; lanes outside [0, ld) load a clamped column (the
        // values are never owned, and the reflected halo or the edge fix-ups
        // supply the ghost columns the owned cells read)
        auto load_row_clamped = [&](int i) {
            const int ii = min(i, y_hi - 1);
            return ld_row4((BUFS ? src : a.in) + static_cast<long long>(ii) * a.ld + col4c);
        };

        // Register rings, indexed by a compile-time phase so the window
        // advances without moves:
        //   level 0: six slots holding rows i-2 .. i+3; at phase ph (mod 6)
        //     row i (newest) is slot ph, i-1 is (ph+5)%6, i-2 is (ph+4)%6,
        //     and row i+3 is loaded into slot (ph+3)%6, three rows ahead;
        //   level j >= 1: three slots; at phase ph (mod 3) the newest row is
        //     slot ph, the one before (ph+2)%3, the oldest (ph+1)%3.
        // The steady-state loop runs six phases per trip (the common period),
        // so every register keeps one role; the general body always runs at
        // phase 0 and rotates afterwards.
        float4 Z[6];
        float4 W[K > 1 ? K - 1 : 1][3];   // W[j-1] is level j, j = 1..K-1
#pragma unroll
        for (int q = 0; q < 3; ++q) Z[q] = load_row(y_lo + q);

        auto win_new = [&](auto PHC, auto JC) -> float4& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[ph % 6];
            else return W[j - 1][ph % 3];
        };
        auto win_mid = [&](auto PHC, auto JC) -> float4& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[(ph + 5) % 6];
            else return W[j - 1][(ph + 2) % 3];
        };
        auto win_old = [&](auto PHC, auto JC) -> float4& {
            constexpr int ph = decltype(PHC)::value, j = decltype(JC)::value;
            if constexpr (j == 0) return Z[(ph + 4) % 6];
            else return W[j - 1][(ph + 1) % 3];
        };

        // One level-j row from level j-1's window.  GEN: any row (edge rows,
        // mirror rows); otherwise a band whose levels touch neither the first
        // two nor the last two rows.  Edge columns are fixed up in both.
        auto level = [&](auto PHC, auto JC, auto GENC, int i) {
            constexpr int j = decltype(JC)::value;
            constexpr bool GEN = decltype(GENC)::value;
            using JP = std::integral_constant<int, j - 1>;
            const int r = i - j;
            // level j is needed on rows [lo, hi): the band and the halo the
            // remaining K - j levels consume (wave-uniform test)
            const int lo = max(ya - (K - j), 0);
            const int hi = min(yb + (K - j), ny);
            if (r < lo || r >= hi) {
                if constexpr (j < K)
                    if (GEN && r == ny) win_new(PHC, JC) = win_mid(PHC, JC);   // row ny := row ny-1
                return;
            }
            const float4 c = win_mid(PHC, JP{});
            const float4 up = win_old(PHC, JP{});
            const float4 dn = win_new(PHC, JP{});
            const float cv[4] = {c.x, c.y, c.z, c.w};
            const float uv[4] = {up.x, up.y, up.z, up.w};
            const float dv[4] = {dn.x, dn.y, dn.z, dn.w};
            // theta' = P(theta) + S + cy*(up + down) + cx*(left + right) with
            // the centre part P(theta) = kc*theta - L(theta) one polynomial,
            // the last term added after the source
            float out4[4], cen4[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float th = cv[e];
                float P;
                if (RAD) P = fmaf(fmaf(fmaf(fmaf(lead, th, a.p3), th, a.p2), th, a.p1), th, a.p0);
                else P = fmaf(lead, th, a.p0);
                out4[e] = fmaf(a.cy, uv[e] + dv[e], P);
                cen4[e] = P;
            }
            const bool own_row = r >= ya && r < yb;
            const bool drow = GEN && BC == LPBF_BC_DIRICHLET && (r == 0 || r == ny - 1);
            // source: the whole row is skipped unless it is near the laser,
            // and only source rows touch the source sum
            const bool src_row = r >= srlo[j - 1] && r <= srhi[j - 1];
            if (src_row) {
                const float uy = static_cast<float>(r) - yc[j - 1];
                // the row's y factor times the source scale
                const float gy = cS[j - 1] * __builtin_amdgcn_exp2f(-(a.ay * uy) * uy);
#pragma unroll
                for (int e = 0; e < 4; ++e) out4[e] = fmaf(gx[j - 1][e], gy, out4[e]);
                // a select, not a nested branch (which would route the
                // wave-uniform condition through a VGPR on every row)
                acc[j - 1].SS = fmaf((own_row && !drow) ? gy : 0.f, gxs[j - 1], acc[j - 1].SS);
                pin(acc[j - 1].SS);
            }
            const float wl = from_left(c.w);    // column left of c.x
            const float wr = from_right(c.x);   // column right of c.w
            const float xm[4] = {wl, c.x, c.y, c.z};
            const float xp[4] = {c.y, c.z, c.w, wr};
#pragma unroll
            for (int e = 0; e < 4; ++e) out4[e] = fmaf(a.cx, xm[e] + xp[e], out4[e]);
            float4 res = make_float4(out4[0], out4[1], out4[2], out4[3]);
            const float4 P4 = make_float4(cen4[0], cen4[1], cen4[2], cen4[3]);
            if (drow) {
                res = make_float4(a.theta_bc, a.theta_bc, a.theta_bc, a.theta_bc);
            } else if (!kReflect && edge) {
                if (BC == LPBF_BC_DIRICHLET) fix_dirichlet(res);
                else fix_mirror(res);
            }
            if (own_row) {
                if (j == K && own_lane)
                    st_row4((BUFS ? dst : a.out) + static_cast<long long>(r) * a.ld + col4, res);
                LevelAcc& A = acc[j - 1];
                float4 mx = res;
                if (ragged && lane == lastl) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (e > laste) elem(mx, e) = kNegInf;
                }
                // running peak of the lane, branch-free: the value, the body
                // row it came from, and that row's four values in LDS (the
                // column is read back once, after the row loop)
                const float m4 = max4(mx);
                const bool up = m4 > A.bv;
                A.bv = up ? m4 : A.bv;
                A.bi = up ? i : A.bi;
                if (up) s_best[j - 1][wave][lane] = mx;
                // molten cells and rows are counted for the wavefront from
                // ballots, so they live in scalar registers
                if (__builtin_amdgcn_ballot_w64(m4 > a.theta_m) & own_mask) {
                    A.molten += __builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(mx.x > a.theta_m) & own_mask) +
                                __builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(mx.y > a.theta_m) & own_mask) +
                                __builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(mx.z > a.theta_m) & own_mask) +
                                __builtin_popcountll(
                                    __builtin_amdgcn_ballot_w64(mx.w > a.theta_m) & own_mask);
                    A.ymin = min(A.ymin, r);
                    A.ymax = r;
                }
                // the loss sum is kc * (sum of the input theta, the previous
                // level's S) - (sum of the centre parts), formed in FP64 from
                // the wavefront sums: the same cells enter both sums.  The
                // per-lane centre sum stays FP32 (relative error ~1e-4 of a
                // step's loss sum, from its rounding at the hot lanes)
                A.S += drow ? 0.0 : static_cast<double>(row_sum(dir_cols(res)));
                A.SP += drow ? 0.f : row_sum(dir_cols(P4));
                // keep each running sum updated here: left alone the compiler
                // defers the serial adds to the end of the unrolled trip and
                // holds every phase's partial live (about 60 extra VGPRs)
                pin(A.S);
                pin(A.SP);
            }
            if constexpr (j < K) {
                win_new(PHC, JC) = res;
                if (GEN && r == 0) win_mid(PHC, JC) = res;   // row -1 := row 0
            }
        };

        using J0 = std::integral_constant<int, 0>;
        auto body = [&](auto PHC, auto GENC, int i) {
            constexpr int ph = decltype(PHC)::value;
            constexpr bool GEN = decltype(GENC)::value;
            Z[(ph + 3) % 6] = GEN ? load_row(i + 3) : load_row_clamped(i + 3);
            float4& v = win_new(PHC, J0{});
            if (!GEN || i < y_hi) {
                if (kReflect) {
                    if (edge) {
                        if (rev) v = make_float4(v.w, v.z, v.y, v.x);
                        if (RAGGED && has_right && pade != 0) fill_right(v);
                        if (RAGGED && nx < 4) {
                            // columns 0..3 hold right ghosts too: lane 0
                            // mirrors lane 1's quad once that is filled
                            const float4 r1 = make_float4(from_right(v.w), from_right(v.z),
                                                          from_right(v.y), from_right(v.x));
                            if (lane == 0) v = r1;
                        }
                    }
                } else if (edge && BC == LPBF_BC_MIRROR) {
                    fix_mirror(v);
                }
                if (i >= ya && i < yb) {
                    const bool drow = GEN && BC == LPBF_BC_DIRICHLET && (i == 0 || i == ny - 1);
                    S0 += drow ? 0.0 : static_cast<double>(row_sum(dir_cols(v)));
                    pin(S0);
                }
            } else if (i == ny) {
                v = win_mid(PHC, J0{});                 // row ny := row ny-1
            }
            if (GEN && i == 0) win_mid(PHC, J0{}) = v;  // row -1 := row 0
            // levels run one after the other: the next level needs this
            // level's row anyway, and interleaving them only raises pressure
            level(PHC, std::integral_constant<int, 1>{}, GENC, i);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (K >= 2) level(PHC, std::integral_constant<int, 2>{}, GENC, i);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (K >= 3) level(PHC, std::integral_constant<int, 3>{}, GENC, i);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (K >= 4) level(PHC, std::integral_constant<int, 4>{}, GENC, i);
        };
        using Gen = std::true_type;
        using Fast = std::false_type;
        auto general = [&](int i) {
            body(std::integral_constant<int, 0>{}, Gen{}, i);
            // back to the phase-0 layout
            const float4 z0 = Z[0];
#pragma unroll
            for (int q = 0; q < 5; ++q) Z[q] = Z[q + 1];
            Z[5] = z0;
#pragma unroll
            for (int q = 0; q + 1 < K; ++q) { W[q][1] = W[q][2]; W[q][2] = W[q][0]; }
        };

        // Fast rows: no level on row 0 or ny-1 (mirror and Dirichlet rows)
        // or next to them.  A band clear of both domain edges runs fast from
        // its first halo row, in six-row trips that may pass i_end: those
        // rows load clamped rows and fail every level's row test.
        const bool clear = ya >= K + 1 && yb + K <= ny - 2;
        const int i_s = clear ? y_lo : max(ya + K, K + 1);
        const int i_e = clear ? i_end + 6 : min(yb + 1, ny - 1);
        int i = y_lo;
        for (; i <= i_end && i < i_s; ++i) general(i);
        // scheduling barriers keep the phases apart: interleaving them buys
        // no latency the two-plus resident waves do not already hide, and
        // costs registers
        // quarter marks of this wavefront's rows, for its issue priority
        const int span = i_end + 1 - y_lo;
        const int q1 = y_lo + span / 4, q2 = y_lo + span / 2, q3 = y_lo + (3 * span) / 4;
        for (; i + 6 <= i_e; i += 6) {
            // a wavefront that is ahead yields issue slots to the ones behind
            if (i >= q3) __builtin_amdgcn_s_setprio(0);
            else if (i >= q2) __builtin_amdgcn_s_setprio(1);
            else if (i >= q1) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(3);
            body(std::integral_constant<int, 0>{}, Fast{}, i);
            __builtin_amdgcn_sched_barrier(0);
            body(std::integral_constant<int, 1>{}, Fast{}, i + 1);
            __builtin_amdgcn_sched_barrier(0);
            body(std::integral_constant<int, 2>{}, Fast{}, i + 2);
            __builtin_amdgcn_sched_barrier(0);
            body(std::integral_constant<int, 3>{}, Fast{}, i + 3);
            __builtin_amdgcn_sched_barrier(0);
            body(std::integral_constant<int, 4>{}, Fast{}, i + 4);
            __builtin_amdgcn_sched_barrier(0);
            body(std::integral_constant<int, 5>{}, Fast{}, i + 5);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (; i <= i_end; ++i) general(i);

        // peak index: level j's best lane value came from row bi - j; the
        // first of that row's four columns holding it
#pragma unroll
        for (int j = 0; j < K; ++j) {
            LevelAcc& A = acc[j];
            if (A.bi != 0x7fffffff) {
                const float4 v = s_best[j][wave][lane];
                const int e = (v.x == A.bv) ? 0 : (v.y == A.bv) ? 1 : (v.z == A.bv) ? 2 : 3;
                A.bi = (A.bi - (j + 1)) * nx + col4 + e;
            }
        }
    }

    // per level: wavefront then workgroup merge -> a.part[j*nwg + wg]
    __shared__ PartialAcc s_pa[kWaves];
    __shared__ double s_sum[4][kWaves];
    __shared__ long long s_m[kWaves];
    const long long N = static_cast<long long>(nx) * ny;
    if (!owner) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            acc[j].bv = kNegInf;
            acc[j].bi = 0x7fffffff;
            acc[j].S = 0.0;
            acc[j].SP = acc[j].SS = 0.f;
        }
        S0 = 0.0;
    }
    double prevS = wave_sum(S0);   // lane 0 holds the wavefront sums
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const LevelAcc& A = acc[j];
        PartialAcc p;
        p.tmax = (A.bi == 0x7fffffff) ? -1.0 : a.Tref_d + static_cast<double>(A.bv);
        p.arg = (A.bi == 0x7fffffff) ? N : static_cast<long long>(A.bi);
        p.ymin = A.ymin;
        p.ymax = A.ymax;
        const PartialAcc w = wave_reduce(p);
        const double wS = wave_sum(A.S);
        const double ws = wave_sum(static_cast<double>(A.SS));
        const double wl = a.kc_d * prevS - wave_sum(static_cast<double>(A.SP));   // dt*loss
        const long long wm = A.molten;   // already the wavefront's count
        if (lane == 0) {
            s_pa[wave] = w;
            s_sum[0][wave] = wS - prevS;   // sum of (theta' - theta) over the owned cells
            s_sum[1][wave] = ws;
            s_sum[2][wave] = wl;
            s_m[wave] = wm;
        }
        prevS = wS;
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

// Built-in band height: about six resident wavefront tiles per SIMD (1024
// SIMDs) so latency is hidden, but never below 8*depth rows, where the
// 2*depth recomputed halo rows would dominate.  lpbf_f32_autotune measures
// instead of guessing.
inline int rows_per_band(int nx, int ny, int depth) {
    const long long nstrips = (nx + kOutCols - 1) / kOutCols;
    const long long bands = (6 * 1024 + nstrips - 1) / nstrips;
    int R = static_cast<int>((ny + bands - 1) / bands);
    const int lo = depth > 1 ? 8 * depth : 16;
    if (R < lo) R = lo;
    if (R > ny) R = ny;
    return R < 1 ? 1 : R;
}

struct Geometry {
    int nstrips, R, ntasks, nwg;
};

// The band height is evened out over the bands it implies: `rows` (or the
// built-in height) sets the band count, ceil(ny / rows), and the bands then
// take ceil(ny / count) rows each, so the last band is not left short while
// every other wavefront carries the full height.
inline Geometry geometry(int nx, int ny, int depth, int rows) {
    Geometry g;
    g.nstrips = (nx + kOutCols - 1) / kOutCols;
    const int R0 = rows > 0 ? std::min(rows, ny) : rows_per_band(nx, ny, depth);
    const int nbands = (ny + R0 - 1) / R0;
    g.R = (ny + nbands - 1) / nbands;
    g.ntasks = g.nstrips * nbands;
    g.nwg = (g.ntasks + kWaves - 1) / kWaves;
    return g;
}

// The FP32 kernels index cells with 32-bit ints and rely on nx, ny for every
// address; shapes outside this range are rejected.
inline bool f32_shape_ok(long long nx, long long ny) {
    return nx >= 2 && ny >= 1 && nx <= (1LL << 28) && ny <= (1LL << 28) &&
           nx * ny < (1LL << 31) - 1;
}

// dt*loss(theta) in K per step as a polynomial in theta = T - T_ref
// (FP64): c[0] + c[1] theta + ... + c[4] theta^4, from
//   cH*(theta + T_ref - Tamb) + cB*(theta + T_ref - Tsub)
//   + cR*((T_ref + theta)^4 - Tamb^4).
inline void loss_poly(const StepConst& k, double T_ref, double c[5]) {
    const double dt = k.dt;
    const double cH = dt * k.hConv * k.inv_h / k.rho_cp;
    const double cB = dt * (k.K / k.dSub) * k.inv_h / k.rho_cp;
    const double cR = k.radiation ? dt * k.eps * 5.670374419e-8 * k.inv_h / k.rho_cp : 0.0;
    const double Tr = T_ref, Ta = k.Tamb;
    c[0] = cH * (Tr - Ta) + cB * (Tr - k.Tsub) + cR * (Tr * Tr * Tr * Tr - Ta * Ta * Ta * Ta);
    c[1] = cH + cB + 4.0 * cR * Tr * Tr * Tr;
    c[2] = 6.0 * cR * Tr * Tr;
    c[3] = 4.0 * cR * Tr;
    c[4] = cR;
}

// Replaces the diffusion numbers and the centre polynomial of `a`.
inline void set_coeffs(F32Args& a, double cx, double cy, const double p[5]) {
    const double kc = 1.0 - 2.0 * (cx + cy);
    a.cx = static_cast<float>(cx);
    a.cy = static_cast<float>(cy);
    a.kc = static_cast<float>(kc);
    a.kc_d = kc;
    a.p0 = static_cast<float>(p[0]);
    a.p1 = static_cast<float>(p[1]);
    a.p2 = static_cast<float>(p[2]);
    a.p3 = static_cast<float>(p[3]);
    a.p4 = static_cast<float>(p[4]);
    a.a0 = static_cast<float>(-p[0]);
    a.a1 = static_cast<float>(kc - p[1]);
    a.a2 = static_cast<float>(-p[2]);
    a.a3 = static_cast<float>(-p[3]);
    a.a4 = static_cast<float>(-p[4]);
}

inline F32Args make_args(const float* in, float* out, int ld, const StepConst& k, double T_ref,
                         int bc, double T_bc, double Tm, int depth, int rows, void* work) {
    F32Args a{};
    const Geometry g = geometry(k.nx, k.ny, depth, rows);
    a.in = in; a.out = out; a.nx = k.nx; a.ny = k.ny; a.ld = ld;
    a.nstrips = g.nstrips; a.R = g.R; a.ntasks = g.ntasks; a.nwg = g.nwg;
    a.bc = bc;
    const double dt = k.dt;
    a.cx = static_cast<float>(k.alpha * dt * k.inv_dx2);
    a.cy = static_cast<float>(k.alpha * dt * k.inv_dy2);
    const double kc = 1.0 - 2.0 * k.alpha * dt * (k.inv_dx2 + k.inv_dy2);
    a.kc = static_cast<float>(kc);
    a.kc_d = kc;
    double L[5];
    loss_poly(k, T_ref, L);
    a.a0 = static_cast<float>(L[0]);
    a.a1 = static_cast<float>(L[1]);
    a.a2 = static_cast<float>(L[2]);
    a.a3 = static_cast<float>(L[3]);
    a.a4 = static_cast<float>(L[4]);
    a.p0 = static_cast<float>(-L[0]);
    a.p1 = static_cast<float>(kc - L[1]);
    a.p2 = static_cast<float>(-L[2]);
    a.p3 = static_cast<float>(-L[3]);
    a.p4 = static_cast<float>(-L[4]);
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
