// SPDX-License-Identifier: 0BSD
// Shared device helpers of the gfx950 step library: error recording, the
// Partial / StepStats merge rules and wavefront (64-lane) reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "lpbf_step_const.h"

namespace lpbf_amd {

constexpr int kWave = 64;
constexpr int kXcds = 8;    // accelerator complex dies of an MI355X

// Workgroup ids are dealt to the XCDs round-robin (id % 8); this maps them so
// that each XCD runs a contiguous range of tiles, which then share halo
// cache lines in that XCD's L2.
__device__ inline int xcd_remap(int bid, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int xcd = bid % kXcds, local = bid / kXcds;
    return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Band heights an autotuner times: a fixed ladder, plus for m = 1..6 rounds
// of `slots` resident wavefronts the height whose band count puts the
// wavefront count just under m whole rounds (a last round that is only
// partly filled leaves SIMDs idle at the end of the launch).  0 stands for
// the built-in height.  Heights below `min_rows` or above ny are dropped.
inline std::vector<int> band_candidates(int ny, long long nstrips, long long slots,
                                        std::initializer_list<int> ladder, int min_rows) {
    std::vector<int> out;
    for (int R : ladder)
        if (R == 0 || (R >= min_rows && R <= ny)) out.push_back(R);
    for (long long m = 1; m <= 6 && nstrips > 0; ++m) {
        const long long bands = std::max(1LL, m * slots / nstrips);
        const int R = static_cast<int>((ny + bands - 1) / bands);
        if (R >= min_rows && R <= ny) out.push_back(R);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

// Records the first error since the last lpbf_get_last_error() query.
void record_error(int err);

// Coefficients that replace a FP32 step launch's own (derived from its
// StepConst): the x and y diffusion numbers and the centre polynomial
// P(theta) = p[0] + p[1] theta + ... + p[4] theta^4, i.e. kc*theta minus
// the loss over the step.  The two-level run's coarse grid takes these so
// that one coarse step reproduces four fine steps' loss (lpbf_multirate.hip).
struct F32Coeffs {
    double cx, cy;
    double p[5];
};

// lpbf_launch_step_f32 (mirror edges, no statistics) with `c` in place of
// the step's own coefficients; the arguments are the caller's to check.
void step_f32_launch_coeffs(const float* theta_in, float* theta_out, int ld, const StepConst& k,
                            double T_ref, const F32Coeffs& c, const LaserTick* d_plan, int step0,
                            int depth, int rows, void* work, hipStream_t stream);

// After a launch: keep the runtime's sticky error visible to the caller
// (reference convention, gpu_step_task.hpp:158) and record it here too.
inline void check_launch() {
    const hipError_t e = hipPeekAtLastError();
    if (e != hipSuccess) record_error(static_cast<int>(e));
}

// Max value with the smallest row-major index on ties, min ymin, max ymax:
// the merge of kernels.cu:229-233 and lpbf_cuda.cu:265-272.  It is exact,
// associative and commutative, so any merge tree gives the serial result.
struct PartialAcc {
    double    tmax;
    long long arg;
    int       ymin, ymax;

    __host__ __device__ static PartialAcc identity(int nx, int ny) {
        return {-1.0, static_cast<long long>(nx) * ny, ny, -1};
    }
    __host__ __device__ void take_max(double v, long long a) {
        if (v > tmax || (v == tmax && a < arg)) { tmax = v; arg = a; }
    }
    __host__ __device__ __forceinline__ void merge(const PartialAcc& o) {
        take_max(o.tmax, o.arg);
        if (o.ymin < ymin) ymin = o.ymin;
        if (o.ymax > ymax) ymax = o.ymax;
    }
};

// Wavefront reduction of a PartialAcc; lane 0 ends with the merged value.
__device__ inline PartialAcc wave_reduce(PartialAcc p) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        PartialAcc o;
        o.tmax = __shfl_down(p.tmax, off);
        o.arg = __shfl_down(p.arg, off);
        o.ymin = __shfl_down(p.ymin, off);
        o.ymax = __shfl_down(p.ymax, off);
        p.merge(o);
    }
    return p;
}

// Fixed-shape tree sum over the wavefront (lane 0 holds the result); the
// order is the same on every launch, so FP64 sums are reproducible.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Lane 0 receives the wavefront maximum.
__device__ inline float wave_max_f(float v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off));
    return v;
}

__device__ inline long long wave_sum(long long v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

// Wavefront reductions on DPP lane moves and lane reads instead of LDS
// permutes: within each 16-lane row, partners by quad_perm [1,0,3,2], then
// [2,3,0,1], then the half-row and row mirrors (both lanes of a pair compute
// a+b from the same two values, so every lane of a row ends with the same
// bits); then the four row results, read from lanes 0, 16, 32, 48, in a
// fixed order.  The result is the same in every lane, and the order does not
// depend on the launch, so the FP64 sums are reproducible.
template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_mov_dpp(v, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ long long dpp_i64(long long v) {
    const int lo = dpp_i32<kCtrl>(static_cast<int>(v));
    const int hi = dpp_i32<kCtrl>(static_cast<int>(v >> 32));
    return (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo);
}
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
    return __longlong_as_double(dpp_i64<kCtrl>(__double_as_longlong(v)));
}
__device__ __forceinline__ long long lane_i64(long long v, int l) {
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(v), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), l);
    return (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo);
}
__device__ __forceinline__ double lane_f64(double v, int l) {
    return __longlong_as_double(lane_i64(__double_as_longlong(v), l));
}

__device__ __forceinline__ double wave_total(double v) {
    v = v + dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v = v + dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v = v + dpp_f64<0x141>(v);   // row_half_mirror
    v = v + dpp_f64<0x140>(v);   // row_mirror
    return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

// Largest value with the smallest index on ties (exact, so any pairing
// gives the serial result).
__device__ __forceinline__ void take_hot(double& t, long long& a, double t2, long long a2) {
    if (t2 > t || (t2 == t && a2 < a)) { t = t2; a = a2; }
}
template <int kCtrl>
__device__ __forceinline__ void hot_step(double& t, long long& a) {
    take_hot(t, a, dpp_f64<kCtrl>(t), dpp_i64<kCtrl>(a));
}
__device__ __forceinline__ void wave_hot(double& t, long long& a) {
    hot_step<0xB1>(t, a);
    hot_step<0x4E>(t, a);
    hot_step<0x141>(t, a);
    hot_step<0x140>(t, a);
    double rt = lane_f64(t, 0);
    long long ra = lane_i64(a, 0);
#pragma unroll
    for (int r = 16; r < kWave; r += 16) take_hot(rt, ra, lane_f64(t, r), lane_i64(a, r));
    t = rt;
    a = ra;
}

__host__ __device__ inline StepStats stats_identity(int nx, int ny) {
    StepStats s;
    s.tmax = -1.0; s.arg = static_cast<long long>(nx) * ny;
    s.ymin = ny; s.ymax = -1; s.molten = 0;
    s.sum_dT = 0.0; s.sum_src = 0.0; s.sum_loss = 0.0;
    return s;
}

// Merge parts[0..n) into *out with one 256-thread workgroup: thread t folds
// t, t+256, ... in order, then a fixed tree, so FP64 sums are reproducible
// and the Partial prefix equals the serial merge.  Thread 0 also returns the
// merged record (the others return an unspecified value).
__device__ inline StepStats block_merge_stats(const StepStats* __restrict__ parts, int n, int nx,
                                              int ny, StepStats* __restrict__ out) {
    constexpr int kWaves = 256 / kWave;
    PartialAcc p = PartialAcc::identity(nx, ny);
    double sd = 0.0, ss = 0.0, sl = 0.0;
    long long m = 0;
    auto fold = [&](const StepStats& q) {
        p.merge(PartialAcc{q.tmax, q.arg, q.ymin, q.ymax});
        m += q.molten; sd += q.sum_dT; ss += q.sum_src; sl += q.sum_loss;
    };
    // four records in flight per thread, folded in the same order as one at
    // a time (the sums do not change)
    int i = threadIdx.x;
    for (; i + 3 * 256 < n; i += 4 * 256) {
        const StepStats q0 = parts[i], q1 = parts[i + 256], q2 = parts[i + 512],
                        q3 = parts[i + 768];
        fold(q0);
        fold(q1);
        fold(q2);
        fold(q3);
    }
    for (; i < n; i += 256) fold(parts[i]);
    __shared__ PartialAcc s_pa[kWaves];
    __shared__ double s_sum[3][kWaves];
    __shared__ long long s_m[kWaves];
    const PartialAcc w = wave_reduce(p);
    const double wd = wave_sum(sd), ws = wave_sum(ss), wl = wave_sum(sl);
    const long long wm = wave_sum(m);
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
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
        for (int i = 1; i < kWaves; ++i) {
            b.merge(s_pa[i]);
            st.molten += s_m[i]; st.sum_dT += s_sum[0][i];
            st.sum_src += s_sum[1][i]; st.sum_loss += s_sum[2][i];
        }
        st.tmax = b.tmax; st.arg = b.arg; st.ymin = b.ymin; st.ymax = b.ymax;
        *out = st;
        return st;
    }
    return StepStats{};
}

// SplitMix64 output function (Steele, Lea & Flood, OOPSLA 2014) applied to
// the counter idx + (seed+1)*golden; returns u in [-0.5, 0.5) with 24 bits.
__host__ __device__ inline double noise_u(unsigned long long seed,
                                          unsigned long long idx) {
    unsigned long long z = idx + (seed + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return static_cast<double>(z >> 40) * (1.0 / 16777216.0) - 0.5;
}

}  // namespace lpbf_amd
