// SPDX-License-Identifier: 0BSD
// Batched FP32 instances (include/lpbf_batch.h).
//
// One launch steps a group of instances of uneven size: the grid is the
// concatenation of every member's workgroups, a table maps a global
// workgroup to its instance, and the instance's arguments (grid, physics,
// buffers, plan, band height) are read from a device array with scalar loads
// (the index is uniform), so each workgroup runs the same tile body as the
// single-field kernel.  The step's statistics of all members are merged by a
// second kernel, one workgroup per (instance, level).  Instances alternate
// between their two buffers; the launch parity picks the direction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <stdexcept>
#include <vector>

#include "lpbf_batch.h"
#include "lpbf_f32_kernel.hpp"

namespace lpbf_amd {
namespace {

struct InstArgs {
    F32Args a;                 // a.in / a.out unused: the buffers are f[parity], f[parity^1]
    float* f[2];
    const LaserTick* plan;
    StepStats* stats;          // per-step records or null
    int wg0;                   // first global workgroup of this instance
    int pad_;
};

// depths 1-3 fit four waves per SIMD (at most 128 VGPRs)
template <int K, int BC, bool RAD>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 4 ? 2 : 4))) void k_step_f32_multi(const InstArgs* __restrict__ inst,
                                                             const int* __restrict__ wg_inst,
                                                             int total_wg, int parity,
                                                             int step0) {
    const int g = xcd_remap(blockIdx.x, total_wg);
    const InstArgs& I = inst[wg_inst[g]];
    step_tile<K, BC, RAD, true>(I.a, I.f[parity], I.f[parity ^ 1], I.plan, step0, g - I.wg0);
}

// Block (i, j): merge level j of instance i into its record for step step0+j.
__global__ __launch_bounds__(kThreads) void k_stats_finalize_multi(
    const InstArgs* __restrict__ inst, int depth, int step0) {
    const int i = blockIdx.x / depth, j = blockIdx.x % depth;
    const InstArgs& I = inst[i];
    if (I.stats == nullptr) return;
    block_merge_stats(I.a.part + static_cast<long long>(j) * I.a.nwg, I.a.nwg, I.a.nx, I.a.ny,
                      I.stats + step0 + j);
}

// Block m: the hot column of job m after `step` (lpbf_launch_hot_cols_f32).
__global__ __launch_bounds__(kThreads) void k_hot_cols_f32(const lpbf_f32_col_job* __restrict__ jobs,
                                                           int q, int step, int slot,
                                                           long long col_pitch) {
    const lpbf_f32_col_job& J = jobs[blockIdx.x];
    if (step >= J.n_steps || !J.cruise[step]) return;
    const int xh = static_cast<int>(J.stats[step].arg % J.nx);
    const float* f = J.theta[q];
    double* col = J.col + static_cast<long long>(slot) * col_pitch;
    for (int y = threadIdx.x; y < J.ny; y += kThreads)
        col[y] = J.T_ref + static_cast<double>(f[static_cast<long long>(y) * J.ld + xh]);
}

template <int K, int BC>
void launch_kb(bool rad, const InstArgs* inst, const int* wg_inst, int total_wg, int parity,
               int step0, hipStream_t s) {
    if (rad)
        hipLaunchKernelGGL((k_step_f32_multi<K, BC, true>), dim3(total_wg), dim3(kThreads), 0, s,
                           inst, wg_inst, total_wg, parity, step0);
    else
        hipLaunchKernelGGL((k_step_f32_multi<K, BC, false>), dim3(total_wg), dim3(kThreads), 0, s,
                           inst, wg_inst, total_wg, parity, step0);
}

template <int K>
void launch_k(int bc, bool rad, const InstArgs* inst, const int* wg_inst, int total_wg,
              int parity, int step0, hipStream_t s) {
    if (bc == LPBF_BC_DIRICHLET)
        launch_kb<K, LPBF_BC_DIRICHLET>(rad, inst, wg_inst, total_wg, parity, step0, s);
    else
        launch_kb<K, LPBF_BC_MIRROR>(rad, inst, wg_inst, total_wg, parity, step0, s);
}

struct Failure {
    int code;
};

void ok(hipError_t e) {
    if (e != hipSuccess) throw Failure{static_cast<int>(e)};
}

// Launch sizes for n steps at depth K: full launches, a shorter last one, and
// an odd count made even by splitting the last multi-step launch (a replayed
// graph must leave every field in the buffer it started in).
std::vector<int> schedule(int n, int K) {
    std::vector<int> d;
    for (int s = 0; s < n; s += K) d.push_back(std::min(K, n - s));
    if (d.size() % 2 == 1) {
        for (size_t q = d.size(); q-- > 0;)
            if (d[q] >= 2) {
                const int x = d[q];
                d[q] = x - 1;
                d.insert(d.begin() + static_cast<long>(q) + 1, 1);
                break;
            }
    }
    return d;
}

}  // namespace

struct Batch {
    struct Class {
        int depth = 0, rows = 0;   // depth 0: not yet set (built-in choice)
    };
    struct Group {
        int K = 1;
        bool rad = true;
        std::vector<int> members;
        InstArgs* d_inst = nullptr;
        int* d_wg_inst = nullptr;
        void* d_work = nullptr;
        int total_wg = 0;
        bool any_stats = false;
        int p = 0;                 // launch parity: members read f[p], write f[p^1]
        std::vector<int> base;     // per member: 1 if f[0] is theta1
    };

    std::vector<lpbf_f32_instance> inst;
    std::vector<int> cur;          // buffer holding each instance's latest field
    int bc = LPBF_BC_MIRROR;
    std::map<int, Class> classes;
    std::vector<Group> groups;
    bool groups_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> graph_parity;   // each group's parity when the graph was captured

    ~Batch() {
        drop_graph();
        drop_groups();
    }

    void drop_graph() {
        if (exec) {
            (void)hipDeviceSynchronize();   // a replay may still be running
            (void)hipGraphExecDestroy(exec);
        }
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_parity.clear();
    }

    static void free_group(Group& g) {
        if (g.d_inst) (void)hipFree(g.d_inst);
        if (g.d_wg_inst) (void)hipFree(g.d_wg_inst);
        if (g.d_work) (void)hipFree(g.d_work);
        g.d_inst = nullptr;
        g.d_wg_inst = nullptr;
        g.d_work = nullptr;
    }

    void drop_groups() {
        for (Group& g : groups) free_group(g);
        groups.clear();
        groups_valid = false;
    }

    int depth_of(int c) const {
        auto it = classes.find(c);
        return (it == classes.end() || it->second.depth == 0) ? 4 : it->second.depth;
    }
    int rows_of(int c) const {
        auto it = classes.find(c);
        return it == classes.end() ? 0 : it->second.rows;
    }
    bool rad_of(int i) const { return inst[static_cast<size_t>(i)].k.radiation; }
    bool plans_cover(long long step0, long long n) const {
        for (auto const& I : inst)
            if (step0 + n > I.n_steps) return false;
        return true;
    }

    // Device arguments for `members` at depth K (rows per member's class, or
    // `rows_override` >= 0 for all).  Each member's f[0] is the buffer that
    // holds its latest field now, so the group starts at parity 0 whatever
    // the members' earlier history.
    Group build(std::vector<int> const& members, int K, bool rad, int rows_override,
                bool with_stats) {
        Group g;
        g.K = K;
        g.rad = rad;
        g.members = members;
        std::vector<InstArgs> h(members.size());
        std::vector<int> wg_inst;
        long long wg = 0;
        for (size_t m = 0; m < members.size(); ++m) {
            const lpbf_f32_instance& I = inst[static_cast<size_t>(members[m])];
            int rows = rows_override >= 0 ? rows_override : rows_of(I.size_class);
            if (rows == 0) rows = rows_per_band(I.k.nx, I.k.ny, K);
            F32Args a = make_args(nullptr, nullptr, I.ld, I.k, I.T_ref, bc, I.T_bc, I.Tm, K, rows,
                                  nullptr);
            const int b0 = cur[static_cast<size_t>(members[m])];
            g.base.push_back(b0);
            h[m].a = a;
            h[m].f[0] = b0 ? I.theta1 : I.theta0;
            h[m].f[1] = b0 ? I.theta0 : I.theta1;
            h[m].plan = I.d_plan;
            h[m].stats = with_stats ? I.d_stats : nullptr;
            h[m].wg0 = static_cast<int>(wg);
            g.any_stats |= h[m].stats != nullptr;
            for (int q = 0; q < a.nwg; ++q) wg_inst.push_back(static_cast<int>(m));
            wg += a.nwg;
        }
        if (wg > (1LL << 30)) throw Failure{LPBF_ERR_INVALID_ARGUMENT};
        g.total_wg = static_cast<int>(wg);
        try {
            ok(hipMalloc(&g.d_work, sizeof(StepStats) * static_cast<size_t>(K) *
                                        static_cast<size_t>(std::max<long long>(wg, 1))));
            for (size_t m = 0; m < members.size(); ++m)
                h[m].a.part = static_cast<StepStats*>(g.d_work) + static_cast<long long>(K) * h[m].wg0;
            ok(hipMalloc(reinterpret_cast<void**>(&g.d_inst), sizeof(InstArgs) * h.size()));
            ok(hipMalloc(reinterpret_cast<void**>(&g.d_wg_inst),
                         sizeof(int) * std::max<size_t>(wg_inst.size(), 1)));
            ok(hipMemcpy(g.d_inst, h.data(), sizeof(InstArgs) * h.size(), hipMemcpyHostToDevice));
            ok(hipMemcpy(g.d_wg_inst, wg_inst.data(), sizeof(int) * wg_inst.size(),
                         hipMemcpyHostToDevice));
        } catch (...) {
            free_group(g);
            throw;
        }
        return g;
    }

    void ensure_groups() {
        if (groups_valid) return;
        drop_graph();
        drop_groups();
        std::map<std::pair<int, bool>, std::vector<int>> by;   // (depth, radiation)
        for (size_t i = 0; i < inst.size(); ++i)
            by[{depth_of(inst[i].size_class), rad_of(static_cast<int>(i))}].push_back(
                static_cast<int>(i));
        for (auto& kv : by)
            groups.push_back(build(kv.second, kv.first.first, kv.first.second, -1, true));
        groups_valid = true;
    }

    int parity(Group const& g) const { return g.p; }
    void set_parity(Group& g, int p) {
        g.p = p;
        for (size_t m = 0; m < g.members.size(); ++m)
            cur[static_cast<size_t>(g.members[m])] = g.base[m] ^ p;
    }
    void flip(Group& g) { set_parity(g, g.p ^ 1); }

    void launch(Group& g, int d, int step0, hipStream_t s) {
        switch (d) {
            case 1: launch_k<1>(bc, g.rad, g.d_inst, g.d_wg_inst, g.total_wg, parity(g), step0, s); break;
            case 2: launch_k<2>(bc, g.rad, g.d_inst, g.d_wg_inst, g.total_wg, parity(g), step0, s); break;
            case 3: launch_k<3>(bc, g.rad, g.d_inst, g.d_wg_inst, g.total_wg, parity(g), step0, s); break;
            default: launch_k<4>(bc, g.rad, g.d_inst, g.d_wg_inst, g.total_wg, parity(g), step0, s); break;
        }
        check_launch();
        if (g.any_stats) {
            hipLaunchKernelGGL(k_stats_finalize_multi, dim3(static_cast<int>(g.members.size()) * d),
                               dim3(kThreads), 0, s, g.d_inst, d, step0);
            check_launch();
        }
        flip(g);
    }

    void enqueue_group(Group& g, int step0, int n, hipStream_t s) {
        int at = step0;
        for (int d : schedule(n, g.K)) {
            launch(g, d, at, s);
            at += d;
        }
    }
};

}  // namespace lpbf_amd

struct lpbf_f32_batch {
    lpbf_amd::Batch b;
};

using lpbf_amd::Batch;
using lpbf_amd::Failure;

namespace {

int fail(int code) {
    lpbf_amd::record_error(code);
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (Failure const& e) {
        return fail(e.code);
    } catch (std::bad_alloc const&) {
        return fail(static_cast<int>(hipErrorOutOfMemory));
    } catch (...) {
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    }
}

bool instance_ok(lpbf_f32_instance const& I) {
    const uintptr_t a0 = reinterpret_cast<uintptr_t>(I.theta0);
    const uintptr_t a1 = reinterpret_cast<uintptr_t>(I.theta1);
    return lpbf_amd::f32_shape_ok(I.k.nx, I.k.ny) && I.theta0 && I.theta1 &&
           I.theta0 != I.theta1 && !(a0 & 15) && !(a1 & 15) && I.ld >= I.k.nx && !(I.ld & 3) &&
           I.d_plan && I.size_class >= 0 && I.k.dt > 0.0 && I.n_steps >= 1;
}

}  // namespace

extern "C" {

int lpbf_f32_batch_create(const lpbf_f32_instance* inst, int n_inst, int bc,
                          lpbf_f32_batch** out) {
    if (!inst || n_inst < 1 || !out || (bc != LPBF_BC_MIRROR && bc != LPBF_BC_DIRICHLET))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    for (int i = 0; i < n_inst; ++i)
        if (!instance_ok(inst[i])) return fail(LPBF_ERR_INVALID_ARGUMENT);
    *out = nullptr;
    return guarded([&] {
        auto* p = new lpbf_f32_batch;
        p->b.inst.assign(inst, inst + n_inst);
        p->b.cur.assign(static_cast<size_t>(n_inst), 0);
        p->b.bc = bc;
        for (int i = 0; i < n_inst; ++i) p->b.classes[inst[i].size_class];
        *out = p;
    });
}

void lpbf_f32_batch_destroy(lpbf_f32_batch* b) {
    delete b;
}

int lpbf_f32_batch_set_tiling(lpbf_f32_batch* b, int size_class, int depth, int rows) {
    if (!b || depth < 1 || depth > lpbf_amd::kMaxDepth || rows < 0 || (rows > 0 && rows < 2 * depth))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    auto it = b->b.classes.find(size_class);
    if (it == b->b.classes.end()) return fail(LPBF_ERR_INVALID_ARGUMENT);
    it->second.depth = depth;
    it->second.rows = rows;
    b->b.groups_valid = false;
    b->b.drop_graph();
    return 0;
}

int lpbf_f32_batch_get_tiling(const lpbf_f32_batch* b, int size_class, int* depth, int* rows,
                              int* n_members) {
    if (!b) return fail(LPBF_ERR_INVALID_ARGUMENT);
    auto it = b->b.classes.find(size_class);
    if (it == b->b.classes.end()) return fail(LPBF_ERR_INVALID_ARGUMENT);
    if (depth) *depth = b->b.depth_of(size_class);
    if (rows) *rows = it->second.rows;
    if (n_members) {
        int n = 0;
        for (auto const& I : b->b.inst) n += I.size_class == size_class;
        *n_members = n;
    }
    return 0;
}

int lpbf_f32_batch_enqueue(lpbf_f32_batch* b, int step0, int n_steps, hipStream_t stream) {
    if (!b || step0 < 0 || n_steps < 0 || !b->b.plans_cover(step0, n_steps))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    return guarded([&] {
        b->b.ensure_groups();
        for (auto& g : b->b.groups) b->b.enqueue_group(g, step0, n_steps, stream);
    });
}

int lpbf_f32_batch_capture(lpbf_f32_batch* b, int step0, int n_steps, hipStream_t stream) {
    if (!b || step0 < 0 || n_steps < 1 || !b->b.plans_cover(step0, n_steps))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    return guarded([&] {
        Batch& B = b->b;
        B.ensure_groups();
        for (auto const& g : B.groups)
            if (lpbf_amd::schedule(n_steps, g.K).size() % 2 == 1)
                throw Failure{LPBF_ERR_INVALID_ARGUMENT};
        B.drop_graph();
        std::vector<int> p0;
        for (auto const& g : B.groups) p0.push_back(g.p);
        auto restore = [&] {   // nothing ran during capture
            for (size_t q = 0; q < B.groups.size(); ++q) B.set_parity(B.groups[q], p0[q]);
        };
        lpbf_amd::ok(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            for (auto& g : B.groups) B.enqueue_group(g, step0, n_steps, stream);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            restore();
            throw;
        }
        restore();   // and an even launch count leaves every field where it started
        lpbf_amd::ok(hipStreamEndCapture(stream, &B.graph));
        lpbf_amd::ok(hipGraphInstantiateWithFlags(&B.exec, B.graph, 0));
        B.graph_parity = p0;
    });
}

int lpbf_f32_batch_launch_graph(lpbf_f32_batch* b, hipStream_t stream) {
    if (!b || !b->b.exec) return fail(LPBF_ERR_INVALID_ARGUMENT);
    // the captured launches read and write fixed buffers: an eager enqueue
    // with an odd launch count since the capture has moved some field to the
    // other buffer, and a replay would then step a stale field
    Batch const& B = b->b;
    if (B.graph_parity.size() != B.groups.size()) return fail(LPBF_ERR_INVALID_ARGUMENT);
    for (size_t q = 0; q < B.groups.size(); ++q)
        if (B.groups[q].p != B.graph_parity[q]) return fail(LPBF_ERR_INVALID_ARGUMENT);
    return guarded([&] { lpbf_amd::ok(hipGraphLaunch(b->b.exec, stream)); });
}

int lpbf_f32_batch_enqueue_launch(lpbf_f32_batch* b, int step0, int n_steps,
                                  hipStream_t stream) {
    if (!b || step0 < 0 || n_steps < 1 || n_steps > lpbf_amd::kMaxDepth ||
        !b->b.plans_cover(step0, n_steps))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    for (auto const& I : b->b.inst)
        if (n_steps > b->b.depth_of(I.size_class)) return fail(LPBF_ERR_INVALID_ARGUMENT);
    return guarded([&] {
        b->b.ensure_groups();
        for (auto& g : b->b.groups) b->b.launch(g, n_steps, step0, stream);
    });
}

int lpbf_launch_hot_cols_f32(const lpbf_f32_col_job* d_jobs, int n_jobs, int q, int step,
                             int slot, long long col_pitch, hipStream_t stream) {
    if (!d_jobs || n_jobs < 0 || (q != 0 && q != 1) || step < 0 || slot < 0 || col_pitch < 0)
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    if (n_jobs == 0) return 0;
    hipLaunchKernelGGL(lpbf_amd::k_hot_cols_f32, dim3(n_jobs), dim3(lpbf_amd::kThreads), 0, stream,
                       d_jobs, q, step, slot, col_pitch);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(static_cast<int>(e));
}

int lpbf_f32_batch_current(const lpbf_f32_batch* b, int instance) {
    if (!b || instance < 0 || instance >= static_cast<int>(b->b.inst.size())) {
        lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
        return -LPBF_ERR_INVALID_ARGUMENT;
    }
    return b->b.cur[static_cast<size_t>(instance)];
}

int lpbf_f32_batch_launches(const lpbf_f32_batch* b, int n_steps, int* n_launches,
                            long long* cells_per_step) {
    if (!b || n_steps < 0) return fail(LPBF_ERR_INVALID_ARGUMENT);
    std::map<std::pair<int, bool>, int> seen;
    int nl = 0;
    long long cells = 0;
    for (size_t i = 0; i < b->b.inst.size(); ++i) {
        auto const& I = b->b.inst[i];
        cells += static_cast<long long>(I.k.nx) * I.k.ny;
        const int K = b->b.depth_of(I.size_class);
        if (seen.emplace(std::make_pair(K, b->b.rad_of(static_cast<int>(i))), 1).second)
            nl += static_cast<int>(lpbf_amd::schedule(n_steps, K).size());
    }
    if (n_launches) *n_launches = nl;
    if (cells_per_step) *cells_per_step = cells;
    return 0;
}

int lpbf_f32_batch_autotune(lpbf_f32_batch* b, int max_depth, hipStream_t stream,
                            double* ms_per_step) {
    if (!b || max_depth < 1 || max_depth > lpbf_amd::kMaxDepth ||
        !b->b.plans_cover(0, max_depth))
        return fail(LPBF_ERR_INVALID_ARGUMENT);
    Batch& B = b->b;
    B.drop_graph();
    std::vector<float*> saved(B.inst.size(), nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double total = 0.0;
    const int rc = guarded([&] {
        // keep every latest field: tuning launches overwrite both buffers
        for (size_t i = 0; i < B.inst.size(); ++i) {
            auto const& I = B.inst[i];
            const size_t bytes = sizeof(float) * static_cast<size_t>(I.ld) * I.k.ny;
            lpbf_amd::ok(hipMalloc(reinterpret_cast<void**>(&saved[i]), bytes));
            lpbf_amd::ok(hipMemcpyAsync(saved[i], B.cur[i] ? I.theta1 : I.theta0, bytes,
                                        hipMemcpyDeviceToDevice, stream));
        }
        lpbf_amd::ok(hipEventCreate(&e0));
        lpbf_amd::ok(hipEventCreate(&e1));
        static const int kRows[] = {0, 16, 24, 32, 40, 48, 64, 96, 128};
        constexpr double kTie = 1.02;   // timings this close count as equal
        const int d_first = max_depth == 1 ? 1 : 2;
        // ms per step of one launch group (two windows of six launches,
        // the faster kept)
        auto time_group = [&](std::vector<int> const& members, int d, bool rad, int rows) {
            Batch::Group g = B.build(members, d, rad, rows, false);
            double t = 1e300;
            try {
                for (int w = 0; w < 2; ++w) {
                    B.launch(g, d, 0, stream);   // warm-up
                    lpbf_amd::ok(hipEventRecord(e0, stream));
                    for (int q = 0; q < 6; ++q) B.launch(g, d, 0, stream);
                    lpbf_amd::ok(hipEventRecord(e1, stream));
                    lpbf_amd::ok(hipEventSynchronize(e1));
                    float ms = 0.f;
                    lpbf_amd::ok(hipEventElapsedTime(&ms, e0, e1));
                    t = std::min(t, ms / (6.0 * d));
                }
            } catch (...) {
                Batch::free_group(g);
                throw;
            }
            Batch::free_group(g);
            return t;
        };
        // the launch groups of the current tiling: classes of equal depth
        // (and radiation switch) share a launch
        auto groups_now = [&] {
            std::map<std::pair<int, bool>, std::vector<int>> sh;
            for (size_t i = 0; i < B.inst.size(); ++i)
                sh[{B.depth_of(B.inst[i].size_class), B.rad_of(static_cast<int>(i))}].push_back(
                    static_cast<int>(i));
            return sh;
        };
        auto schedule_ms = [&] {
            double t = 0.0;
            for (auto const& sg : groups_now())
                t += time_group(sg.second, sg.first.first, sg.first.second, -1);
            return t;
        };

        // 1. each class alone: per depth, the fastest band height (within
        //    timing noise of it, the tallest: fewer halo rows and less
        //    traffic, which tells most when classes share the device)
        std::map<int, std::map<int, std::pair<double, int>>> at;   // class -> depth -> (ms, rows)
        std::map<int, int> min_ny;
        for (auto& kv : B.classes) {
            std::map<bool, std::vector<int>> mem;   // by radiation switch
            int mny = 1 << 30;
            for (size_t i = 0; i < B.inst.size(); ++i)
                if (B.inst[i].size_class == kv.first) {
                    mem[B.rad_of(static_cast<int>(i))].push_back(static_cast<int>(i));
                    mny = std::min(mny, B.inst[i].k.ny);
                }
            min_ny[kv.first] = mny;
            for (int d = d_first; d <= max_depth; ++d) {
                std::vector<std::pair<double, int>> timed;
                for (int R : kRows) {
                    if (R > 0 && (R < 2 * d || R > mny)) continue;
                    double ms = 0.0;
                    for (auto const& m : mem) ms += time_group(m.second, d, m.first, R);
                    timed.push_back({ms, R});
                }
                std::pair<double, int> b{1e300, 0};
                for (auto const& t : timed)
                    if (t.first < b.first) b = t;
                for (auto const& t : timed)
                    if (t.first <= b.first * kTie && t.second > b.second && b.second != 0)
                        b.second = t.second;
                at[kv.first][d] = b;
            }
        }
        // 2. schedules: every class at its own best depth, or every class at
        //    one depth (one shared launch per step group instead of several
        //    small ones); each timed as its launches, the fastest kept
        //    (within timing noise, the one with fewer launch groups)
        std::vector<std::map<int, std::pair<int, int>>> cands;   // class -> (depth, rows)
        {
            std::map<int, std::pair<int, int>> own;
            for (auto const& c : at) {
                int bd = d_first;
                for (auto const& dv : c.second)
                    if (dv.second.first < c.second.at(bd).first) bd = dv.first;
                own[c.first] = {bd, c.second.at(bd).second};
            }
            cands.push_back(own);
            for (int d = d_first; d <= max_depth; ++d) {
                std::map<int, std::pair<int, int>> one;
                for (auto const& c : at) one[c.first] = {d, c.second.at(d).second};
                if (one != own) cands.push_back(one);
            }
        }
        double best_ms = 1e300;
        size_t best_groups = 0;
        std::map<int, std::pair<int, int>> pick;
        for (auto const& cand : cands) {
            for (auto const& c : cand) {
                B.classes[c.first].depth = c.second.first;
                B.classes[c.first].rows = c.second.second;
            }
            const double t = schedule_ms();
            const size_t ng = groups_now().size();
            const bool better = best_ms == 1e300 || t < best_ms / kTie ||
                                (t <= best_ms * kTie && ng < best_groups);
            if (better) { best_ms = t; best_groups = ng; pick = cand; }
        }
        for (auto const& c : pick) {
            B.classes[c.first].depth = c.second.first;
            B.classes[c.first].rows = c.second.second;
        }
        // 3. in each shared launch, the band heights picked alone need not be
        //    best (wave rounds, the halo traffic of the whole launch): one
        //    pass of coordinate descent, a class's band height at a time,
        //    timing the shared launch itself
        for (auto const& sg : groups_now()) {
            const int d = sg.first.first;
            std::map<int, int> cls;   // classes in this launch -> smallest ny
            for (int i : sg.second) cls[B.inst[static_cast<size_t>(i)].size_class] = 0;
            if (cls.size() < 2) continue;
            double best = time_group(sg.second, d, sg.first.second, -1);
            for (auto const& cm : cls) {
                Batch::Class& cl = B.classes[cm.first];
                const int keep = cl.rows;
                std::vector<std::pair<double, int>> timed{{best, keep}};
                for (int R : kRows) {
                    if (R == keep || (R > 0 && (R < 2 * d || R > min_ny[cm.first]))) continue;
                    cl.rows = R;
                    timed.push_back({time_group(sg.second, d, sg.first.second, -1), R});
                }
                int br = keep;
                for (auto const& t : timed)
                    if (t.first < best) { best = t.first; br = t.second; }
                for (auto const& t : timed)
                    if (t.first <= best * kTie && t.second > br && br != 0) br = t.second;
                cl.rows = br;
            }
        }
        total = schedule_ms();
        const int e = lpbf_get_last_error();
        if (e) throw Failure{e};
    });
    // put every field back into theta0 (tuning left the buffers scrambled)
    int rc2 = 0;
    for (size_t i = 0; i < B.inst.size(); ++i) {
        if (!saved[i]) continue;
        auto const& I = B.inst[i];
        const size_t bytes = sizeof(float) * static_cast<size_t>(I.ld) * I.k.ny;
        if (hipMemcpyAsync(I.theta0, saved[i], bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            rc2 = LPBF_ERR_INVALID_ARGUMENT;
        B.cur[i] = 0;
    }
    if (hipStreamSynchronize(stream) != hipSuccess && !rc2) rc2 = static_cast<int>(hipErrorUnknown);
    for (float* p : saved)
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    B.groups_valid = false;
    if (rc) return rc;
    if (rc2) return fail(rc2);
    if (ms_per_step) *ms_per_step = total;
    return 0;
}

}  // extern "C"
