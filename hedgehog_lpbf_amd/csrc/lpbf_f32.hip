// SPDX-License-Identifier: 0BSD
// FP32 performance path for gfx950, single field: the temporally blocked
// step kernel (tile body in lpbf_f32_kernel.hpp), its statistics merge,
// noise, conversions, the hot column, the autotuner and the C ABI.
#include <mutex>

#include "lpbf_f32_kernel.hpp"

namespace lpbf_amd {
namespace {

// Occupancy targets: depth 4 with radiation fits three waves per SIMD (at
// most 168 VGPRs), without it two; depths 1-3 fit four (at most 128).
template <int K, int BC, bool RAD, bool RAGGED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 4 ? (RAD ? 3 : 2) : 4))) void k_step_f32(F32Args a,
                                                       const LaserTick* __restrict__ plan,
                                                       int step0) {
    step_tile<K, BC, RAD, false, RAGGED>(a, nullptr, nullptr, plan, step0,
                                         xcd_remap(blockIdx.x, a.nwg));
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

// Column x_hot = glob->arg % nx of T = T_ref + theta, as FP64.
__global__ void k_extract_col_f32(const float* __restrict__ th, int ld, int nx, int ny,
                                  double Tref, const StepStats* __restrict__ glob,
                                  double* __restrict__ col) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    const int x_hot = static_cast<int>(glob->arg % nx);
    col[y] = Tref + static_cast<double>(th[static_cast<long long>(y) * ld + x_hot]);
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

template <int K, bool RAD, bool RAGGED>
void launch_krg(const F32Args& a, const LaserTick* plan, int step0, hipStream_t s) {
    if (a.bc == LPBF_BC_DIRICHLET)
        hipLaunchKernelGGL((k_step_f32<K, LPBF_BC_DIRICHLET, RAD, RAGGED>), dim3(a.nwg),
                           dim3(kThreads), 0, s, a, plan, step0);
    else
        hipLaunchKernelGGL((k_step_f32<K, LPBF_BC_MIRROR, RAD, RAGGED>), dim3(a.nwg),
                           dim3(kThreads), 0, s, a, plan, step0);
}

// A plate whose width is a multiple of 4 has no lane straddling column nx.
template <int K, bool RAD>
void launch_kr(const F32Args& a, const LaserTick* plan, int step0, hipStream_t s) {
    if (a.nx % 4 == 0) launch_krg<K, RAD, false>(a, plan, step0, s);
    else launch_krg<K, RAD, true>(a, plan, step0, s);
}

template <int K>
void launch_k(const F32Args& a, const LaserTick* plan, int step0, hipStream_t s) {
    if (a.a4 != 0.f) launch_kr<K, true>(a, plan, step0, s);
    else launch_kr<K, false>(a, plan, step0, s);
}

bool step_f32_args_ok(const float* theta_in, const float* theta_out, int ld, const StepConst* k,
                      int bc, const LaserTick* d_plan, int step0, int depth, int rows) {
    return k && f32_shape_ok(k->nx, k->ny) && theta_in && theta_out && theta_in != theta_out &&
           ld >= k->nx && !(ld & 3) && !(reinterpret_cast<uintptr_t>(theta_in) & 15) &&
           !(reinterpret_cast<uintptr_t>(theta_out) & 15) && d_plan && step0 >= 0 &&
           depth >= 1 && depth <= kMaxDepth && rows >= 0 &&
           (bc == LPBF_BC_MIRROR || bc == LPBF_BC_DIRICHLET);
}

// The step launch and, with d_stats, its statistics merge; `work` holds the
// per-workgroup records (every record is written before the merge reads it).
void step_f32_launch(const float* theta_in, float* theta_out, int ld, const StepConst* k,
                     double T_ref, int bc, double T_bc, double Tm, const LaserTick* d_plan,
                     int step0, int depth, int rows, StepStats* d_stats, void* work,
                     hipStream_t stream) {
    const F32Args a = make_args(theta_in, theta_out, ld, *k, T_ref, bc, T_bc, Tm, depth, rows,
                                work);
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

// Workspace of lpbf_launch_step_f32_noalloc: one array per device, for the
// callers that cannot allocate device memory.  In host code `scratch` names
// a host-side stand-in; its device address comes from hipGetSymbolAddress,
// looked up once per device (with that device current) and kept here.
__device__ __align__(256) unsigned char scratch[LPBF_F32_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// The current device's scratch address, or nullptr (error recorded).
void* scratch_of_current_device() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        record_error(static_cast<int>(e));
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            record_error(e != hipSuccess ? static_cast<int>(e) : LPBF_ERR_INVALID_ARGUMENT);
            return nullptr;
        }
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

}  // namespace

void step_f32_launch_coeffs(const float* theta_in, float* theta_out, int ld, const StepConst& k,
                            double T_ref, const F32Coeffs& c, const LaserTick* d_plan, int step0,
                            int depth, int rows, void* work, hipStream_t stream) {
    F32Args a = make_args(theta_in, theta_out, ld, k, T_ref, LPBF_BC_MIRROR, T_ref, T_ref + 1e9,
                          depth, rows, work);
    set_coeffs(a, c.cx, c.cy, c.p);
    switch (depth) {
        case 1: launch_k<1>(a, d_plan, step0, stream); break;
        case 2: launch_k<2>(a, d_plan, step0, stream); break;
        case 3: launch_k<3>(a, d_plan, step0, stream); break;
        default: launch_k<4>(a, d_plan, step0, stream); break;
    }
    check_launch();
}

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
    if (!step_f32_args_ok(theta_in, theta_out, ld, k, bc, d_plan, step0, depth, rows) ||
        !d_work) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    step_f32_launch(theta_in, theta_out, ld, k, T_ref, bc, T_bc, Tm, d_plan, step0, depth, rows,
                    d_stats, d_work, stream);
}

void lpbf_launch_step_f32_noalloc(const float* theta_in, float* theta_out, int ld,
                                  const StepConst* k, double T_ref, int bc, double T_bc,
                                  double Tm, const LaserTick* d_plan, int step0, int depth,
                                  int rows, StepStats* d_stats, hipStream_t stream) {
    if (!step_f32_args_ok(theta_in, theta_out, ld, k, bc, d_plan, step0, depth, rows)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const Geometry g = geometry(k->nx, k->ny, depth, rows);
    const size_t need = sizeof(StepStats) * static_cast<size_t>(g.nwg) * static_cast<size_t>(depth);
    if (need > LPBF_F32_NOALLOC_SCRATCH_BYTES) {
        record_error(LPBF_ERR_SCRATCH_LIMIT);
        return;
    }
    void* work = scratch_of_current_device();
    if (!work) return;
    step_f32_launch(theta_in, theta_out, ld, k, T_ref, bc, T_bc, Tm, d_plan, step0, depth, rows,
                    d_stats, work, stream);
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

int lpbf_f32_autotune(const StepConst* k, double T_ref, int bc, double T_bc, double Tm,
                      const LaserTick* d_plan, int n_plan, int max_depth, hipStream_t stream,
                      int* best_depth, int* best_rows, double* best_ms_per_step) {
    if (!k || !f32_shape_ok(k->nx, k->ny) || !d_plan || max_depth < 1 || max_depth > kMaxDepth ||
        n_plan < max_depth || !best_depth || !best_rows ||
        (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET)) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return LPBF_ERR_INVALID_ARGUMENT;
    }
    const int nx = k->nx, ny = k->ny, ld = (nx + 3) & ~3;
    // three resident waves per SIMD at depth 4, 1024 SIMDs
    const std::vector<int> rows_c = band_candidates(
        ny, (nx + kOutCols - 1) / kOutCols, 3 * 1024, {0, 16, 24, 32, 40, 48, 56, 64, 96, 128, 192},
        1);
    size_t work_bytes = 0;
    for (int d = 1; d <= max_depth; ++d)
        for (int R : rows_c) {
            const size_t b = lpbf_f32_workspace_bytes(nx, ny, d, R);
            if (b > work_bytes) work_bytes = b;
        }
    float* a = nullptr;
    float* b = nullptr;
    void* work = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t field = sizeof(float) * static_cast<size_t>(ld) * ny;
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&a), field);
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&b), field);
    if (err == hipSuccess) err = hipMalloc(&work, work_bytes);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    double best = 1e300;
    int bd = 1, br = 0;
    if (err == hipSuccess) {
        lpbf_launch_noise_f32(a, nx, ny, ld, 1, 0.0, 0.5, stream);
        const int reps = 10;   // long enough that launch-to-launch noise does not pick the tiling
        const int d_first = (max_depth == 1) ? 1 : 2;
        for (int depth = d_first; depth <= max_depth && err == hipSuccess; ++depth) {
            for (int R : rows_c) {
                if (R > 0 && R < 2 * depth) continue;
                lpbf_launch_step_f32(a, b, ld, k, T_ref, bc, T_bc, Tm, d_plan, 0, depth, R, nullptr,
                                     work, stream);   // warm-up
                err = hipEventRecord(e0, stream);
                if (err != hipSuccess) break;
                for (int q = 0; q < reps; ++q)
                    lpbf_launch_step_f32((q & 1) ? a : b, (q & 1) ? b : a, ld, k, T_ref, bc, T_bc,
                                         Tm, d_plan, 0, depth, R, nullptr, work, stream);
                err = hipEventRecord(e1, stream);
                if (err == hipSuccess) err = hipEventSynchronize(e1);
                float ms = 0.f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                if (err != hipSuccess) break;
                const double per_step = ms / (reps * depth);
                if (per_step < best) { best = per_step; bd = depth; br = R; }
            }
        }
    }
    // release everything; the first error (if any) is the one reported
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (work) (void)hipFree(work);
    if (err != hipSuccess) {
        record_error(static_cast<int>(err));
        return static_cast<int>(err);
    }
    *best_depth = bd;
    *best_rows = br;
    if (best_ms_per_step) *best_ms_per_step = best;
    return 0;
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

void lpbf_launch_extract_col_f32(const float* theta, int ld, int nx, int ny, double T_ref,
                                 const StepStats* glob, double* out_col, hipStream_t stream) {
    if (!theta || !glob || !out_col || !f32_shape_ok(nx, ny) || ld < nx) {
        record_error(LPBF_ERR_INVALID_ARGUMENT);
        return;
    }
    const int bs = 128;
    hipLaunchKernelGGL(k_extract_col_f32, dim3((ny + bs - 1) / bs), dim3(bs), 0, stream, theta, ld,
                       nx, ny, T_ref, glob, out_col);
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
