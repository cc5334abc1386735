// SPDX-License-Identifier: 0BSD
// lpbf::DeviceRun and its C ABI (include/lpbf_run.h).
//
// Per batch, on the run's stream and without host synchronisation: step
// launches with fused statistics, the per-step merge into the device ring,
// the hot column on cruise steps, the field copy on dump steps; then the
// rings come back with one copy into pinned memory and an event.  The host
// queues the next batch, waits for the event and finalises the metrics in
// step order (reference include/gpu_step_task.hpp:135-239 and
// MetricsTask::execute, :375-404).
#include "lpbf/device_run.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "lpbf_host_util.hpp"
#include "lpbf_kernels.h"
#include "lpbf_run.h"

namespace lpbf {

std::mutex& setup_mutex() {
    static std::mutex m;
    return m;
}

namespace {

void hip_ok(hipError_t e, const char* what) { hip_check(e, "DeviceRun", what); }
void lib_ok(const char* what) { lib_check("DeviceRun", what); }

hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

DeviceRun::DeviceRun(Domain dom, StepConst k, int phys, double Tm, std::vector<Tick> plan,
                     double Tamb, RunOptions opt)
    : dom_(dom), k_(k), phys_(phys), Tm_(Tm), plan_(std::move(plan)), Tamb_(Tamb), opt_(opt),
      nx_(dom.nx), ny_(dom.ny), met_(dom, Tm) {
    if (k_.nx != nx_ || k_.ny != ny_ || nx_ < 2 || ny_ < 1)
        throw std::invalid_argument("DeviceRun: StepConst grid does not match the domain");
    if (opt_.precision != 64 && opt_.precision != 32)
        throw std::invalid_argument("DeviceRun: precision must be 64 or 32");
    if (opt_.bc != LPBF_BC_MIRROR && opt_.bc != LPBF_BC_DIRICHLET)
        throw std::invalid_argument("DeviceRun: unknown boundary condition");
    if (phys_ != LPBF_PHYS_BASELINE && phys_ != LPBF_PHYS_EXTENDED)
        throw std::invalid_argument("DeviceRun: unknown physics");
    if (opt_.batch < 1) opt_.batch = 1;
    if (opt_.precision == 32 && (opt_.depth < 1 || opt_.depth > 4))
        throw std::invalid_argument("DeviceRun: depth must be 1..4");
    // the FP32 extended physics steps one step per launch
    if (opt_.precision == 32 && phys_ == LPBF_PHYS_EXTENDED) opt_.depth = 1;
    if (opt_.rows < 0) throw std::invalid_argument("DeviceRun: rows must be >= 0");
    if (opt_.f64_depth == 0) opt_.f64_depth = 1;
    if (opt_.precision == 64 && (opt_.f64_depth < 1 || opt_.f64_depth > 4))
        throw std::invalid_argument("DeviceRun: f64_depth must be 1..4");
    // the blocked FP64 launches replace the graph batches
    if (opt_.precision == 64 && opt_.f64_depth > 1) opt_.graph = false;
    T_ref_ = opt_.T_ref < 0.0 ? Tamb_ : opt_.T_ref;

    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    std::lock_guard<std::mutex> lock(setup_mutex());
    try {
        hipStream_t s;
        hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        stream_ = s;
        if (opt_.precision == 64) {
            hip_ok(hipMalloc(&d_a_, N * sizeof(double)), "field");
            hip_ok(hipMalloc(&d_b_, N * sizeof(double)), "field");
            nblk_ = lpbf_step_ex_nblocks(nx_, ny_);
            if (nblk_ <= 0) throw std::invalid_argument("DeviceRun: grid too large");
            std::size_t blk = static_cast<std::size_t>(nblk_) * sizeof(StepStats);
            if (opt_.f64_depth > 1) {
                const std::size_t W =
                    lpbf_f64_blocked_workspace_bytes(nx_, ny_, opt_.f64_depth, opt_.rows);
                if (W == 0) throw std::invalid_argument("DeviceRun: grid outside the blocked range");
                blk = std::max(blk, W);
                std::vector<LaserTick> lt = to_laser(plan_);
                if (!lt.empty()) {
                    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_plan_),
                                     lt.size() * sizeof(LaserTick)), "plan");
                    hip_ok(hipMemcpyAsync(d_plan_, lt.data(), lt.size() * sizeof(LaserTick),
                                          hipMemcpyHostToDevice, S(stream_)), "plan upload");
                    hip_ok(hipStreamSynchronize(S(stream_)), "plan upload");
                }
            }
            hip_ok(hipMalloc(&d_blk_, blk), "block records");
            buf0_ = d_a_;
            if (opt_.graph && opt_.batch % 2 == 0 && !plan_.empty()) {
                std::vector<LaserTick> lt = to_laser(plan_);
                hip_ok(hipMalloc(reinterpret_cast<void**>(&d_plan_), lt.size() * sizeof(LaserTick)),
                       "plan");
                hip_ok(hipMemcpyAsync(d_plan_, lt.data(), lt.size() * sizeof(LaserTick),
                                      hipMemcpyHostToDevice, S(stream_)), "plan upload");
                hip_ok(hipStreamSynchronize(S(stream_)), "plan upload");
                std::vector<unsigned char> cr(plan_.size());
                for (std::size_t i = 0; i < plan_.size(); ++i) cr[i] = cruise(plan_[i]) ? 1 : 0;
                hip_ok(hipMalloc(reinterpret_cast<void**>(&d_cruise_), cr.size()), "cruise flags");
                hip_ok(hipMemcpyAsync(d_cruise_, cr.data(), cr.size(), hipMemcpyHostToDevice,
                                      S(stream_)), "cruise upload");
                hip_ok(hipStreamSynchronize(S(stream_)), "cruise upload");
                hip_ok(hipMalloc(reinterpret_cast<void**>(&d_bticks_), B * sizeof(LaserTick)),
                       "batch ticks");
                hip_ok(hipMalloc(reinterpret_cast<void**>(&d_bcruise_), B), "batch cruise flags");
                // the batch launches keep two steps' block records
                const int pb = 2 * lpbf_step_ex_batch_nblocks(nx_, ny_);
                if (pb > nblk_) {
                    hip_ok(hipFree(d_blk_), "block records");
                    d_blk_ = nullptr;
                    hip_ok(hipMalloc(&d_blk_, static_cast<std::size_t>(pb) * sizeof(StepStats)),
                           "block records");
                }
            }
        } else {
            ld_ = (nx_ + 3) & ~3;
            const std::size_t F = static_cast<std::size_t>(ld_) * ny_ * sizeof(float);
            hip_ok(hipMalloc(&d_a_, F), "field");
            hip_ok(hipMalloc(&d_b_, F), "field");
            const std::size_t W = phys_ == LPBF_PHYS_EXTENDED
                                      ? lpbf_f32_ext_workspace_bytes(nx_, ny_)
                                      : lpbf_f32_workspace_bytes(nx_, ny_, opt_.depth, opt_.rows);
            if (W == 0) throw std::invalid_argument("DeviceRun: grid outside the FP32 range");
            hip_ok(hipMalloc(&d_blk_, W), "workspace");
            std::vector<LaserTick> lt = to_laser(plan_);
            if (!lt.empty()) {
                hip_ok(hipMalloc(reinterpret_cast<void**>(&d_plan_), lt.size() * sizeof(LaserTick)),
                       "plan");
                hip_ok(hipMemcpyAsync(d_plan_, lt.data(), lt.size() * sizeof(LaserTick),
                                      hipMemcpyHostToDevice, S(stream_)), "plan upload");
                hip_ok(hipStreamSynchronize(S(stream_)), "plan upload");
            }
        }
        hip_ok(hipMalloc(reinterpret_cast<void**>(&d_ring_), B * sizeof(StepStats)), "ring");
        hip_ok(hipMalloc(reinterpret_cast<void**>(&d_cols_), B * ny_ * sizeof(double)), "columns");
        // two halves: batch b+1 is queued before the host reads batch b's
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&h_ring_), 2 * B * sizeof(StepStats)),
               "pinned ring");
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&h_cols_), 2 * B * ny_ * sizeof(double)),
               "pinned columns");
        for (int q = 0; q < 2; ++q) {
            hipEvent_t ev;
            hip_ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
            done_ev_[q] = ev;
        }
        if (opt_.keep_records) rec_.reserve(plan_.size());
        std::vector<double> init(N, Tamb_);
        upload(init.data());
    } catch (...) {
        release_locked();
        throw;
    }
}

DeviceRun::~DeviceRun() { release(); }

void DeviceRun::release() noexcept {
    std::lock_guard<std::mutex> lock(setup_mutex());
    release_locked();
}

// Teardown cannot report failures; a sticky device error stays visible to
// the next HIP call of the caller.
void DeviceRun::release_locked() noexcept {
    if (stream_) (void)hipStreamSynchronize(S(stream_));
    for (int q = 0; q < 2; ++q) {
        if (gexec_[q]) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(gexec_[q]));
        if (graph_[q]) (void)hipGraphDestroy(static_cast<hipGraph_t>(graph_[q]));
        gexec_[q] = graph_[q] = nullptr;
    }
    if (d_bticks_) (void)hipFree(d_bticks_);
    if (d_bcruise_) (void)hipFree(d_bcruise_);
    if (d_cruise_) (void)hipFree(d_cruise_);
    d_bticks_ = nullptr;
    d_bcruise_ = nullptr;
    d_cruise_ = nullptr;
    for (void* p : {d_a_, d_b_, d_blk_, static_cast<void*>(d_plan_), static_cast<void*>(d_ring_),
                    static_cast<void*>(d_cols_), static_cast<void*>(d_snap_)})
        if (p) (void)hipFree(p);
    if (h_ring_) (void)hipHostFree(h_ring_);
    if (h_cols_) (void)hipHostFree(h_cols_);
    for (void*& ev : done_ev_) {
        if (ev) (void)hipEventDestroy(static_cast<hipEvent_t>(ev));
        ev = nullptr;
    }
    if (stream_) (void)hipStreamDestroy(S(stream_));
    d_a_ = d_b_ = d_blk_ = nullptr;
    d_plan_ = nullptr; d_ring_ = nullptr; d_cols_ = nullptr; d_snap_ = nullptr;
    h_ring_ = nullptr; h_cols_ = nullptr; stream_ = nullptr;
}

void DeviceRun::upload(const double* T) {
    if (!T) throw std::invalid_argument("DeviceRun::upload: null field");
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    if (opt_.precision == 64) {
        hip_ok(hipMemcpyAsync(d_a_, T, N * sizeof(double), hipMemcpyHostToDevice, S(stream_)),
               "upload");
    } else {
        if (!d_snap_) hip_ok(hipMalloc(reinterpret_cast<void**>(&d_snap_), N * sizeof(double)),
                             "staging");
        hip_ok(hipMemcpyAsync(d_snap_, T, N * sizeof(double), hipMemcpyHostToDevice, S(stream_)),
               "upload");
        lpbf_launch_T_to_theta(d_snap_, nx_, ny_, T_ref_, static_cast<float*>(d_a_), ld_,
                               S(stream_));
    }
    hip_ok(hipStreamSynchronize(S(stream_)), "upload sync");
    lib_ok("upload");
}

void DeviceRun::noise(std::uint64_t seed, double base, double amp) {
    if (opt_.precision == 64)
        lpbf_launch_noise_f64(static_cast<double*>(d_a_),
                              static_cast<long long>(nx_) * ny_, seed, base, amp, S(stream_));
    else
        lpbf_launch_noise_f32(static_cast<float*>(d_a_), nx_, ny_, ld_, seed, base - T_ref_, amp,
                              S(stream_));
    hip_ok(hipStreamSynchronize(S(stream_)), "noise sync");
    lib_ok("noise");
}

void DeviceRun::download(double* T) const {
    if (!T) throw std::invalid_argument("DeviceRun::download: null field");
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    if (opt_.precision == 64) {
        hip_ok(hipMemcpyAsync(T, d_a_, N * sizeof(double), hipMemcpyDeviceToHost, S(stream_)),
               "download");
    } else {
        auto self = const_cast<DeviceRun*>(this);
        if (!self->d_snap_)
            hip_ok(hipMalloc(reinterpret_cast<void**>(&self->d_snap_), N * sizeof(double)),
                   "staging");
        lpbf_launch_theta_to_T(static_cast<const float*>(d_a_), ld_, nx_, ny_, T_ref_, d_snap_,
                               S(stream_));
        hip_ok(hipMemcpyAsync(T, d_snap_, N * sizeof(double), hipMemcpyDeviceToHost, S(stream_)),
               "download");
    }
    hip_ok(hipStreamSynchronize(S(stream_)), "download sync");
    lib_ok("download");
}

// Batches end at multiples of `batch`, so full ones can replay the graph.
// They are pipelined: batch b+1 is queued (launches, ring and column copies
// into the other pinned half, an event) before the host waits for batch b
// and finalises its metrics, so the GPU does not wait for the host.  The
// device ring and columns are reused in stream order.
int DeviceRun::advance(int n) {
    const int start = done_;
    const int target = static_cast<int>(std::min<long long>(
        static_cast<long long>(plan_.size()), static_cast<long long>(done_) + std::max(n, 0)));
    int queued = done_, slot = 0;
    Pending cur, nxt;
    bool have = false;
    while (have || queued < target) {
        bool have_n = false;
        if (queued < target) {
            const int bn = std::min(opt_.batch - queued % opt_.batch, target - queued);
            nxt = enqueue(queued, bn, slot);
            queued += bn;
            slot ^= 1;
            have_n = true;
        }
        if (have) complete(cur);
        cur = std::move(nxt);
        have = have_n;
    }
    return done_ - start;
}

StepRecord RunMetrics::finish(Tick const& tk, StepStats const& st, const double* col) {
    // MetricsTask::execute (gpu_step_task.hpp:375-404), same expressions.
    const int ny = dom_.ny;
    const double dy = dom_.dy();
    const double W_um = (st.ymax < 0) ? 0.0 : (st.ymax - st.ymin + 1) * dy * 1e6;
    double W_sub_um = 0.0;
    if (st.tmax > Tm_ && col) {
        int a = -1, b = -1;
        for (int y = 0; y < ny; ++y)
            if (col[y] > Tm_) {
                if (a < 0) a = y;
                b = y;
            }
        if (a >= 0) {
            double y_lo = a * dy, y_hi = b * dy;
            if (a > 0) y_lo = ((a - 1) + (Tm_ - col[a - 1]) / (col[a] - col[a - 1])) * dy;
            if (b < ny - 1) y_hi = (b + (Tm_ - col[b]) / (col[b + 1] - col[b])) * dy;
            W_sub_um = (y_hi - y_lo) * 1e6;
        }
    }
    if (st.tmax > peak_T_) peak_T_ = st.tmax;
    if (W_um > peak_W_) peak_W_ = W_um;
    if (cruise(tk) && W_sub_um > steady_) steady_ = W_sub_um;
    StepRecord r;
    r.step = tk.step;
    r.has_col = col != nullptr;
    r.stats = st;
    r.W_um = W_um;
    r.W_sub_um = W_sub_um;
    return r;
}

void DeviceRun::finish_step(Tick const& tk, StepStats const& st, const double* col) {
    last_ = met_.finish(tk, st, col);
    if (opt_.keep_records) rec_.push_back(last_);
}

// One batch of opt_.batch steps (even) as kernel nodes, one per step j,
// plus a flush: node j computes step j with fused block statistics while its
// extra workgroup merges step j-1's into ring[j-1] and copies that step's hot
// column on cruise steps (the flush finishes the last step).  Laser ticks and
// cruise flags are read from the batch copies, refreshed before each replay,
// so no node depends on a device counter.  `parity` names the physical
// buffer holding the current field; an even count leaves it there.
void DeviceRun::capture(int parity) {
    const hipStream_t s = S(stream_);
    const int B = opt_.batch;
    std::lock_guard<std::mutex> lock(setup_mutex());
    double* cur = static_cast<double*>(parity == 0 ? buf0_ : (buf0_ == d_a_ ? d_b_ : d_a_));
    double* nxt = static_cast<double*>(cur == d_a_ ? d_b_ : d_a_);
    hipGraph_t g = nullptr;
    hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int i = 0; i < B; ++i) {
        lpbf_launch_step_ex_batch(cur, nxt, &k_, phys_, opt_.bc, opt_.T_bc, Tm_, d_bticks_, i,
                                  d_bcruise_, static_cast<StepStats*>(d_blk_), d_ring_, d_cols_, s);
        std::swap(cur, nxt);
    }
    lpbf_launch_step_ex_batch_flush(cur, nx_, ny_, B, d_bcruise_,
                                    static_cast<const StepStats*>(d_blk_), d_ring_, d_cols_, s);
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        hip_ok(e, "end capture");
    }
    graph_[parity] = g;
    lib_ok("captured launches");
    hipGraphExec_t x = nullptr;
    hip_ok(hipGraphInstantiateWithFlags(&x, g, 0), "instantiate");
    gexec_[parity] = x;
}

bool DeviceRun::graph_batch(int s0, int bn) {
    if (opt_.precision != 64 || !d_bticks_ || bn != opt_.batch || s0 % opt_.batch != 0) return false;
    if (snap_cb_)
        for (int i = 0; i < bn; ++i)
            if (plan_[static_cast<std::size_t>(s0 + i)].dump) return false;
    const int parity = (d_a_ == buf0_) ? 0 : 1;
    if (!gexec_[parity]) capture(parity);
    hip_ok(hipMemcpyAsync(d_bticks_, d_plan_ + s0, static_cast<std::size_t>(bn) * sizeof(LaserTick),
                          hipMemcpyDeviceToDevice, S(stream_)), "batch ticks");
    hip_ok(hipMemcpyAsync(d_bcruise_, d_cruise_ + s0, static_cast<std::size_t>(bn),
                          hipMemcpyDeviceToDevice, S(stream_)), "batch cruise flags");
    hip_ok(hipGraphLaunch(static_cast<hipGraphExec_t>(gexec_[parity]), S(stream_)),
           "graph launch");
    ++graph_launches_;
    return true;
}

DeviceRun::Pending DeviceRun::enqueue(int s0, int bn, int slot) {
    const hipStream_t s = S(stream_);
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    Pending p;
    p.s0 = s0;
    p.bn = bn;
    p.slot = slot;
    std::vector<char>& has_col = p.has_col;
    has_col.assign(static_cast<std::size_t>(bn), 0);
    std::vector<Snap>& pend = p.snaps;

    if (graph_batch(s0, bn)) {
        for (int i = 0; i < bn; ++i)
            has_col[static_cast<std::size_t>(i)] = cruise(plan_[static_cast<std::size_t>(s0 + i)]);
    } else if (opt_.precision == 64 && opt_.f64_depth > 1) {
        int i = 0;
        while (i < bn) {
            // a launch covers up to f64_depth steps and ends at a cruise step
            // (the hot column) or a dump step (the snapshot)
            int L = 1;
            while (L < opt_.f64_depth && i + L < bn) {
                const Tick& t = plan_[static_cast<std::size_t>(s0 + i + L - 1)];
                if (cruise(t) || (t.dump && snap_cb_)) break;
                ++L;
            }
            lpbf_launch_step_f64_blocked(static_cast<const double*>(d_a_),
                                         static_cast<double*>(d_b_), &k_, phys_, opt_.bc,
                                         opt_.T_bc, Tm_, d_plan_, s0 + i, L, opt_.rows,
                                         d_ring_ + i, d_blk_, s);
            std::swap(d_a_, d_b_);
            const int j = i + L - 1;
            const Tick& tk = plan_[static_cast<std::size_t>(s0 + j)];
            if (cruise(tk)) {
                lpbf_launch_extract_col(static_cast<const double*>(d_a_), nx_, ny_,
                                        reinterpret_cast<const Partial*>(d_ring_ + j),
                                        d_cols_ + static_cast<std::size_t>(j) * ny_, s);
                has_col[static_cast<std::size_t>(j)] = 1;
            }
            if (tk.dump && snap_cb_) {
                pend.push_back({tk.step, std::vector<double>(N)});
                hip_ok(hipMemcpyAsync(pend.back().field.data(), d_a_, N * sizeof(double),
                                      hipMemcpyDeviceToHost, s), "snapshot");
            }
            i += L;
        }
    } else if (opt_.precision == 64) {
        for (int i = 0; i < bn; ++i) {
            const Tick& tk = plan_[static_cast<std::size_t>(s0 + i)];
            k_.x_l = tk.x_l; k_.y_l = tk.y_l; k_.q_scale = tk.q_scale;
            lpbf_launch_step_ex(static_cast<const double*>(d_a_), static_cast<double*>(d_b_), &k_,
                                phys_, opt_.bc, opt_.T_bc, Tm_, static_cast<StepStats*>(d_blk_), s);
            std::swap(d_a_, d_b_);
            lpbf_launch_stats_finalize(static_cast<const StepStats*>(d_blk_), nblk_, nx_, ny_,
                                       d_ring_ + i, s);
            if (cruise(tk)) {
                lpbf_launch_extract_col(static_cast<const double*>(d_a_), nx_, ny_,
                                        reinterpret_cast<const Partial*>(d_ring_ + i),
                                        d_cols_ + static_cast<std::size_t>(i) * ny_, s);
                has_col[static_cast<std::size_t>(i)] = 1;
            }
            if (tk.dump && snap_cb_) {
                pend.push_back({tk.step, std::vector<double>(N)});
                hip_ok(hipMemcpyAsync(pend.back().field.data(), d_a_, N * sizeof(double),
                                      hipMemcpyDeviceToHost, s), "snapshot");
            }
        }
    } else {
        int i = 0;
        while (i < bn) {
            // a launch covers up to `depth` steps and ends at a dump step (and,
            // with cruise_cols, at a cruise step)
            int L = 1;
            while (L < opt_.depth && i + L < bn) {
                const Tick& t = plan_[static_cast<std::size_t>(s0 + i + L - 1)];
                if (t.dump || (opt_.cruise_cols && cruise(t))) break;
                ++L;
            }
            if (phys_ == LPBF_PHYS_EXTENDED)
                lpbf_launch_step_f32_ext(static_cast<const float*>(d_a_), static_cast<float*>(d_b_),
                                         ld_, &k_, T_ref_, opt_.bc, opt_.T_bc, Tm_, d_plan_, s0 + i,
                                         d_ring_ + i, d_blk_, s);
            else
                lpbf_launch_step_f32(static_cast<const float*>(d_a_), static_cast<float*>(d_b_),
                                     ld_, &k_, T_ref_, opt_.bc, opt_.T_bc, Tm_, d_plan_, s0 + i, L,
                                     opt_.rows, d_ring_ + i, d_blk_, s);
            std::swap(d_a_, d_b_);
            const int j = i + L - 1;
            const Tick& tk = plan_[static_cast<std::size_t>(s0 + j)];
            if (cruise(tk)) {
                lpbf_launch_extract_col_f32(static_cast<const float*>(d_a_), ld_, nx_, ny_, T_ref_,
                                            d_ring_ + j, d_cols_ + static_cast<std::size_t>(j) * ny_,
                                            s);
                has_col[static_cast<std::size_t>(j)] = 1;
            }
            if (tk.dump && snap_cb_) {
                if (!d_snap_)
                    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_snap_), N * sizeof(double)),
                           "staging");
                lpbf_launch_theta_to_T(static_cast<const float*>(d_a_), ld_, nx_, ny_, T_ref_,
                                       d_snap_, s);
                pend.push_back({tk.step, std::vector<double>(N)});
                hip_ok(hipMemcpyAsync(pend.back().field.data(), d_snap_, N * sizeof(double),
                                      hipMemcpyDeviceToHost, s), "snapshot");
            }
            i += L;
        }
    }
    lib_ok("step launches");
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    hip_ok(hipMemcpyAsync(h_ring_ + slot * B, d_ring_, static_cast<std::size_t>(bn) * sizeof(StepStats),
                          hipMemcpyDeviceToHost, s), "ring copy");
    if (std::find(has_col.begin(), has_col.end(), 1) != has_col.end())
        hip_ok(hipMemcpyAsync(h_cols_ + slot * B * ny_, d_cols_,
                              static_cast<std::size_t>(bn) * ny_ * sizeof(double),
                              hipMemcpyDeviceToHost, s), "column copy");
    hip_ok(hipEventRecord(static_cast<hipEvent_t>(done_ev_[slot]), s), "batch event");
    return p;
}

void DeviceRun::complete(Pending& p) {
    hip_ok(hipEventSynchronize(static_cast<hipEvent_t>(done_ev_[p.slot])), "batch sync");
    lib_ok("batch");
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    const StepStats* hr = h_ring_ + p.slot * B;
    const double* hc = h_cols_ + p.slot * B * ny_;
    for (int i = 0; i < p.bn; ++i)
        finish_step(plan_[static_cast<std::size_t>(p.s0 + i)], hr[i],
                    p.has_col[static_cast<std::size_t>(i)]
                        ? hc + static_cast<std::size_t>(i) * ny_ : nullptr);
    done_ += p.bn;
    for (auto& q : p.snaps) snap_cb_(q.step, std::move(q.field));
}

}  // namespace lpbf

// ---- C ABI -------------------------------------------------------------------

struct lpbf_run {
    lpbf::DeviceRun run;
    template <class... A>
    explicit lpbf_run(A&&... a) : run(std::forward<A>(a)...) {}
};

using lpbf::guarded;

extern "C" {

int lpbf_run_create(const StepConst* k, int phys, double Tm, double Lx, double Ly, double Tamb,
                    const LaserTick* plan, const double* x_csv, const unsigned char* wrapped,
                    const unsigned char* dump, int n_steps, const lpbf_run_options* opt,
                    lpbf_run** out) {
    if (!k || !out || n_steps < 0 || (n_steps > 0 && !plan) || k->nx < 2 || k->ny < 2) {
        lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
        return LPBF_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return guarded([&] {
        lpbf::Domain dom;
        dom.nx = k->nx;
        dom.ny = k->ny;
        dom.Lx = Lx;
        dom.Ly = Ly;
        std::vector<lpbf::Tick> ticks(static_cast<std::size_t>(n_steps));
        for (int s = 0; s < n_steps; ++s) {
            lpbf::Tick& t = ticks[static_cast<std::size_t>(s)];
            t.step = s;
            t.t = s * k->dt;
            t.x_l = plan[s].x_l;
            t.y_l = plan[s].y_l;
            t.q_scale = plan[s].q_scale;
            t.x_laser_csv = x_csv ? x_csv[s] : -1.0;
            t.wrapped = x_csv ? (wrapped ? wrapped[s] != 0 : false) : true;
            t.dump = dump ? dump[s] != 0 : false;
        }
        lpbf::RunOptions o;
        if (opt) {
            o.precision = opt->precision;
            o.bc = opt->bc;
            o.T_bc = opt->T_bc;
            o.batch = opt->batch;
            o.depth = opt->depth;
            o.rows = opt->rows;
            o.T_ref = opt->T_ref;
            o.keep_records = opt->keep_records != 0;
            o.graph = opt->graph != 0;
            o.f64_depth = opt->f64_depth;
            o.cruise_cols = opt->cruise_cols != 0;
        }
        *out = new lpbf_run(dom, *k, phys, Tm, std::move(ticks), Tamb, o);
    });
}

void lpbf_run_destroy(lpbf_run* run) { delete run; }

int lpbf_run_upload(lpbf_run* run, const double* T) {
    if (!run || !T) { lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT); return LPBF_ERR_INVALID_ARGUMENT; }
    return guarded([&] { run->run.upload(T); });
}

int lpbf_run_noise(lpbf_run* run, unsigned long long seed, double base, double amp) {
    if (!run) { lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT); return LPBF_ERR_INVALID_ARGUMENT; }
    return guarded([&] { run->run.noise(seed, base, amp); });
}

int lpbf_run_advance(lpbf_run* run, int n_steps) {
    if (!run || n_steps < 0) {
        lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
        return -LPBF_ERR_INVALID_ARGUMENT;
    }
    int ran = 0;
    const int rc = guarded([&] { ran = run->run.advance(n_steps); });
    return rc ? -rc : ran;
}

int lpbf_run_download(lpbf_run* run, double* T) {
    if (!run || !T) { lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT); return LPBF_ERR_INVALID_ARGUMENT; }
    return guarded([&] { run->run.download(T); });
}

int lpbf_run_summary(const lpbf_run* run, double* out3, int* steps_done) {
    if (!run || !out3) { lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT); return LPBF_ERR_INVALID_ARGUMENT; }
    out3[0] = run->run.peak_T();
    out3[1] = run->run.peak_W();
    out3[2] = run->run.steady_W_sub();
    if (steps_done) *steps_done = run->run.steps_done();
    return 0;
}

int lpbf_run_graph_launches(const lpbf_run* run) {
    if (!run) { lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT); return -LPBF_ERR_INVALID_ARGUMENT; }
    return run->run.graph_launches();
}

int lpbf_run_records(const lpbf_run* run, int first, int n, lpbf_step_record* out) {
    if (!run || !out || first < 0 || n < 0 ||
        static_cast<std::size_t>(first) + static_cast<std::size_t>(n) > run->run.records().size()) {
        lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
        return LPBF_ERR_INVALID_ARGUMENT;
    }
    auto const& r = run->run.records();
    for (int i = 0; i < n; ++i) {
        lpbf::StepRecord const& s = r[static_cast<std::size_t>(first + i)];
        out[i].stats = s.stats;
        out[i].W_um = s.W_um;
        out[i].W_sub_um = s.W_sub_um;
        out[i].step = s.step;
        out[i].has_col = s.has_col ? 1 : 0;
    }
    return 0;
}

}  // extern "C"
