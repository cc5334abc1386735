// SPDX-License-Identifier: 0BSD
// lpbf::Ensemble and its C ABI (include/lpbf_run.h, lpbf_ensemble_*).
//
// Device layout per run m: two FP64 fields, the laser plan and cruise flags
// of the whole run, 2 x nblk block records; shared by all runs: the ring
// [m][batch] of merged per-step statistics and the hot columns
// [batch][m][ny] (step-major, so the cruise steps of a batch are one
// contiguous copy), the batch start step, and the lpbf_ensemble_run table
// the launches read.  Batch step j is one lpbf_launch_step_ensemble over
// every run; the reference counterpart is a loop of run_one_row calls
// (src/gpu/hh_gpu_main.cpp:170-198, 234-290).
#include "lpbf/ensemble.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <utility>

#include "lpbf_batch.h"
#include "lpbf_host_util.hpp"
#include "lpbf_kernels.h"
#include "lpbf_run.h"

namespace lpbf {

namespace {

void hip_ok(hipError_t e, const char* what) { hip_check(e, "Ensemble", what); }
void lib_ok(const char* what) { lib_check("Ensemble", what); }

hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }
lpbf_f32_batch* FB(void* b) { return static_cast<lpbf_f32_batch*>(b); }

void rc_ok(int rc, const char* what) {
    if (rc != 0) lib_check("Ensemble", what);
    if (rc != 0) throw std::runtime_error(std::string("Ensemble: ") + what + " failed");
}

}  // namespace

Ensemble::Ensemble(Domain dom, int phys, double Tm, std::vector<EnsembleMember> runs,
                   EnsembleOptions opt)
    : dom_(dom), phys_(phys), Tm_(Tm), runs_(std::move(runs)), opt_(opt), nx_(dom.nx),
      ny_(dom.ny) {
    const int n = static_cast<int>(runs_.size());
    if (n < 1 || n > 65535) throw std::invalid_argument("Ensemble: 1..65535 runs");
    if (nx_ < 2 || ny_ < 1) throw std::invalid_argument("Ensemble: grid too small");
    if (phys_ != LPBF_PHYS_BASELINE && phys_ != LPBF_PHYS_EXTENDED)
        throw std::invalid_argument("Ensemble: unknown physics");
    if (opt_.bc != LPBF_BC_MIRROR && opt_.bc != LPBF_BC_DIRICHLET)
        throw std::invalid_argument("Ensemble: unknown boundary condition");
    if (opt_.batch < 1) throw std::invalid_argument("Ensemble: batch must be >= 1");
    if (opt_.precision != 64 && opt_.precision != 32)
        throw std::invalid_argument("Ensemble: precision must be 64 or 32");
    for (EnsembleMember const& r : runs_) {
        if (r.k.nx != nx_ || r.k.ny != ny_)
            throw std::invalid_argument("Ensemble: a run's StepConst grid does not match the domain");
        if (r.plan.size() > static_cast<std::size_t>(1 << 30))
            throw std::invalid_argument("Ensemble: plan too long");
        max_steps_ = std::max(max_steps_, static_cast<int>(r.plan.size()));
    }
    if (opt_.precision == 32) {
        if (phys_ != LPBF_PHYS_BASELINE)
            throw std::invalid_argument("Ensemble: the FP32 path runs baseline physics only");
        if (opt_.depth < 1 || opt_.depth > 4)
            throw std::invalid_argument("Ensemble: depth must be 1..4");
        if (opt_.rows < 0 || (opt_.rows > 0 && opt_.rows < 2 * opt_.depth))
            throw std::invalid_argument("Ensemble: rows must be 0 or >= 2*depth");
        opt_.graph = false;
    } else {
        if (opt_.rows < 0 || opt_.rows > 1024)
            throw std::invalid_argument("Ensemble: rows must be 0..1024");
        if (opt_.rows == 0) opt_.rows = lpbf_step_ensemble_rows(nx_, ny_, n);
        nblk_ = lpbf_step_ensemble_nblocks(nx_, ny_, opt_.rows);
        if (nblk_ <= 0) throw std::invalid_argument("Ensemble: grid too large");
        if (static_cast<long long>(nblk_ + 1) * n > (1LL << 31) - 1)
            throw std::invalid_argument("Ensemble: too many runs for one launch");
    }
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    const std::size_t nn = static_cast<std::size_t>(n);
    if (static_cast<long long>(n) * ny_ > (1LL << 31) - 1)
        throw std::invalid_argument("Ensemble: column pitch too large");

    met_.assign(nn, RunMetrics(dom_, Tm_));
    rec_.resize(nn);
    if (opt_.keep_records)
        for (std::size_t m = 0; m < nn; ++m) rec_[m].reserve(runs_[m].plan.size());

    std::lock_guard<std::mutex> lock(setup_mutex());
    try {
        auto dmalloc = [&](std::size_t bytes, const char* what) { return dalloc(bytes, what); };
        hipStream_t s;
        hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        stream_ = s;
        d_ring_ = static_cast<StepStats*>(dmalloc(nn * B * sizeof(StepStats), "ring"));
        d_cols_ = static_cast<double*>(dmalloc(nn * B * ny_ * sizeof(double), "columns"));
        d_s0_ = static_cast<int*>(dmalloc(sizeof(int), "batch start"));
        // two of each host-side buffer: batch b+1 is queued before the host
        // reads batch b's
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&h_ring_), 2 * nn * B * sizeof(StepStats)),
               "pinned ring");
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&h_cols_),
                             2 * nn * B * ny_ * sizeof(double)), "pinned columns");
        hip_ok(hipHostMalloc(reinterpret_cast<void**>(&h_s0_), 2 * sizeof(int)),
               "pinned batch start");
        for (int q = 0; q < 2; ++q) {
            hipEvent_t ev;
            hip_ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
            done_ev_[q] = ev;
        }

        if (opt_.precision == 32) {
            setup_f32();
            return;
        }
        std::vector<lpbf_ensemble_run> table(nn);
        std::vector<double> init(N);
        for (std::size_t m = 0; m < nn; ++m) {
            EnsembleMember const& r = runs_[m];
            const std::size_t ns = r.plan.size();
            lpbf_ensemble_run& e = table[m];
            e.k = r.k;
            e.buf[0] = static_cast<double*>(dmalloc(N * sizeof(double), "field"));
            e.buf[1] = static_cast<double*>(dmalloc(N * sizeof(double), "field"));
            buf_.push_back(e.buf[0]);
            buf_.push_back(e.buf[1]);
            const std::vector<LaserTick> lt = to_laser(r.plan);
            std::vector<unsigned char> cr(ns);
            for (std::size_t i = 0; i < ns; ++i) cr[i] = met_[m].cruise(r.plan[i]) ? 1 : 0;
            LaserTick* d_plan = static_cast<LaserTick*>(dmalloc(ns * sizeof(LaserTick), "plan"));
            unsigned char* d_cr = static_cast<unsigned char*>(dmalloc(ns, "cruise flags"));
            if (ns) {
                hip_ok(hipMemcpy(d_plan, lt.data(), ns * sizeof(LaserTick), hipMemcpyHostToDevice),
                       "plan upload");
                hip_ok(hipMemcpy(d_cr, cr.data(), ns, hipMemcpyHostToDevice), "cruise upload");
            }
            e.plan = d_plan;
            e.cruise = d_cr;
            e.rec = static_cast<StepStats*>(dmalloc(2 * static_cast<std::size_t>(nblk_) *
                                                        sizeof(StepStats), "block records"));
            e.ring = d_ring_ + m * B;
            e.cols = d_cols_ + m * ny_;
            e.n_steps = static_cast<int>(ns);
            e.col_pitch = n * ny_;
            std::fill(init.begin(), init.end(), r.Tamb);
            hip_ok(hipMemcpy(e.buf[0], init.data(), N * sizeof(double), hipMemcpyHostToDevice),
                   "initial field");
        }
        d_runs_ = dmalloc(nn * sizeof(lpbf_ensemble_run), "run table");
        hip_ok(hipMemcpy(d_runs_, table.data(), nn * sizeof(lpbf_ensemble_run),
                         hipMemcpyHostToDevice), "run table upload");
    } catch (...) {
        release();
        throw;
    }
}

Ensemble::~Ensemble() {
    std::lock_guard<std::mutex> lock(setup_mutex());
    release();
}

void* Ensemble::dalloc(std::size_t bytes, const char* what) {
    void* p = nullptr;
    hip_ok(hipMalloc(&p, bytes ? bytes : 1), what);
    dev_.push_back(p);
    return p;
}

// Precision 32: per run two theta fields, its plan, cruise flags and
// max_steps per-step records; the batch of all runs is built by
// rebuild_f32.  Called from the constructor with the set-up lock held.
void Ensemble::setup_f32() {
    const std::size_t nn = runs_.size();
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    ld_ = (nx_ + 3) & ~3;
    const std::size_t F = static_cast<std::size_t>(ld_) * ny_ * sizeof(float);
    const hipStream_t s = S(stream_);
    d_stage_ = static_cast<double*>(dalloc(N * sizeof(double), "staging"));
    d_stats_ = static_cast<StepStats*>(
        dalloc(nn * static_cast<std::size_t>(std::max(max_steps_, 1)) * sizeof(StepStats),
               "per-step records"));
    d_jobs_ = dalloc(nn * sizeof(lpbf_f32_col_job), "column jobs");
    std::vector<double> init(N);
    cruise_.resize(nn);
    fcur_.assign(nn, 0);
    for (std::size_t m = 0; m < nn; ++m) {
        EnsembleMember const& r = runs_[m];
        const std::size_t ns = r.plan.size();
        tref_.push_back(opt_.T_ref < 0.0 ? r.Tamb : opt_.T_ref);
        for (int q = 0; q < 2; ++q) {
            float* f = static_cast<float*>(dalloc(F, "field"));
            hip_ok(hipMemsetAsync(f, 0, F, s), "field clear");
            fbuf_.push_back(f);
        }
        const std::vector<LaserTick> lt = to_laser(r.plan);
        cruise_[m].resize(ns);
        for (std::size_t i = 0; i < ns; ++i) cruise_[m][i] = met_[m].cruise(r.plan[i]) ? 1 : 0;
        d_plan_.push_back(static_cast<LaserTick*>(dalloc(ns * sizeof(LaserTick), "plan")));
        d_cruise_.push_back(static_cast<unsigned char*>(dalloc(ns, "cruise flags")));
        if (ns) {
            hip_ok(hipMemcpyAsync(d_plan_.back(), lt.data(), ns * sizeof(LaserTick),
                                  hipMemcpyHostToDevice, s), "plan upload");
            hip_ok(hipMemcpyAsync(d_cruise_.back(), cruise_[m].data(), ns, hipMemcpyHostToDevice,
                                  s), "cruise upload");
        }
        std::fill(init.begin(), init.end(), r.Tamb);
        hip_ok(hipMemcpyAsync(d_stage_, init.data(), N * sizeof(double), hipMemcpyHostToDevice, s),
               "initial field");
        lpbf_launch_T_to_theta(d_stage_, nx_, ny_, tref_[m], fbuf_[2 * m], ld_, s);
        // the host buffers are reused for the next run
        hip_ok(hipStreamSynchronize(s), "initial field");
        lib_ok("initial field");
    }
    for (std::size_t m = 0; m < nn; ++m)
        if (!runs_[m].plan.empty()) active_.push_back(static_cast<int>(m));
    rebuild_f32(0);
}

// Buffer (0/1) of run m's pair holding its latest field: the launches since
// the last rebuild moved every active run's field once each.
int Ensemble::latest_f32(int m) const {
    const int c = fcur_[static_cast<std::size_t>(m)];
    return std::find(active_.begin(), active_.end(), m) != active_.end() ? c ^ (flaunch_ & 1) : c;
}

// (Re)builds the batch over the runs still running at step `next` (the
// first step not yet queued); the queued work is finished first, so the
// job table and the batch's device arguments are rewritten while idle.
// Called with the set-up lock held.
void Ensemble::rebuild_f32(int next) {
    fresh_ = true;
    const hipStream_t s = S(stream_);
    hip_ok(hipStreamSynchronize(s), "rebuild sync");
    for (int m : active_) fcur_[static_cast<std::size_t>(m)] ^= flaunch_ & 1;
    flaunch_ = 0;
    std::vector<int> keep;
    for (int m : active_)
        if (n_steps(m) > next) keep.push_back(m);
    active_ = keep;
    if (fbatch_) lpbf_f32_batch_destroy(FB(fbatch_));
    fbatch_ = nullptr;
    if (active_.empty()) return;
    const std::size_t na = active_.size();
    std::vector<lpbf_f32_instance> inst(na);
    std::vector<lpbf_f32_col_job> jobs(na);
    for (std::size_t a = 0; a < na; ++a) {
        const int m = active_[a];
        const std::size_t um = static_cast<std::size_t>(m);
        const int c = fcur_[um];
        lpbf_f32_instance& I = inst[a];
        I = lpbf_f32_instance{};
        I.k = runs_[um].k;
        I.theta0 = fbuf_[2 * um + static_cast<std::size_t>(c)];
        I.theta1 = fbuf_[2 * um + static_cast<std::size_t>(c ^ 1)];
        I.ld = ld_;
        I.size_class = 0;
        I.T_ref = tref_[um];
        I.T_bc = opt_.T_bc;
        I.Tm = Tm_;
        I.d_plan = d_plan_[um];
        I.d_stats = d_stats_ + um * static_cast<std::size_t>(max_steps_);
        I.n_steps = n_steps(m);
        lpbf_f32_col_job& J = jobs[a];
        J = lpbf_f32_col_job{};
        J.theta[0] = I.theta0;
        J.theta[1] = I.theta1;
        J.stats = I.d_stats;
        J.cruise = d_cruise_[um];
        J.col = d_cols_ + um * static_cast<std::size_t>(ny_);
        J.T_ref = I.T_ref;
        J.ld = ld_;
        J.nx = nx_;
        J.ny = ny_;
        J.n_steps = I.n_steps;
    }
    lpbf_f32_batch* b = nullptr;
    rc_ok(lpbf_f32_batch_create(inst.data(), static_cast<int>(na), opt_.bc, &b), "batch create");
    fbatch_ = b;
    rc_ok(lpbf_f32_batch_set_tiling(b, 0, opt_.depth, opt_.rows), "batch tiling");
    hip_ok(hipMemcpy(d_jobs_, jobs.data(), na * sizeof(lpbf_f32_col_job), hipMemcpyHostToDevice),
           "column jobs");
}

// Teardown cannot report failures; a sticky device error stays visible to
// the next HIP call of the caller.  Called with the set-up lock held.
void Ensemble::release() noexcept {
    if (stream_) (void)hipStreamSynchronize(S(stream_));
    if (fbatch_) lpbf_f32_batch_destroy(FB(fbatch_));
    fbatch_ = nullptr;
    fbuf_.clear();
    d_stats_ = nullptr; d_jobs_ = nullptr; d_stage_ = nullptr;
    if (gexec_) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(gexec_));
    if (graph_) (void)hipGraphDestroy(static_cast<hipGraph_t>(graph_));
    gexec_ = graph_ = nullptr;
    for (void*& ev : done_ev_) {
        if (ev) (void)hipEventDestroy(static_cast<hipEvent_t>(ev));
        ev = nullptr;
    }
    for (void* p : dev_) (void)hipFree(p);
    dev_.clear();
    buf_.clear();
    if (h_ring_) (void)hipHostFree(h_ring_);
    if (h_cols_) (void)hipHostFree(h_cols_);
    if (h_s0_) (void)hipHostFree(h_s0_);
    if (stream_) (void)hipStreamDestroy(S(stream_));
    h_ring_ = nullptr; h_cols_ = nullptr; h_s0_ = nullptr; stream_ = nullptr;
    d_runs_ = nullptr; d_ring_ = nullptr; d_cols_ = nullptr; d_s0_ = nullptr;
}

void Ensemble::upload(int m, const double* T) {
    if (!T || m < 0 || m >= size()) throw std::invalid_argument("Ensemble::upload: bad run or field");
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    const std::size_t um = static_cast<std::size_t>(m);
    if (opt_.precision == 32) {
        hip_ok(hipMemcpyAsync(d_stage_, T, N * sizeof(double), hipMemcpyHostToDevice, S(stream_)),
               "upload");
        lpbf_launch_T_to_theta(d_stage_, nx_, ny_, tref_[um],
                               fbuf_[2 * um + static_cast<std::size_t>(latest_f32(m))], ld_,
                               S(stream_));
    } else {
        const int cur = std::min(done_, n_steps(m)) & 1;
        hip_ok(hipMemcpyAsync(buf_[2 * um + cur], T, N * sizeof(double), hipMemcpyHostToDevice,
                              S(stream_)), "upload");
    }
    hip_ok(hipStreamSynchronize(S(stream_)), "upload sync");
    lib_ok("upload");
}

void Ensemble::download(int m, double* T) const {
    if (!T || m < 0 || m >= size())
        throw std::invalid_argument("Ensemble::download: bad run or field");
    const std::size_t N = static_cast<std::size_t>(nx_) * ny_;
    const std::size_t um = static_cast<std::size_t>(m);
    if (opt_.precision == 32) {
        lpbf_launch_theta_to_T(fbuf_[2 * um + static_cast<std::size_t>(latest_f32(m))], ld_, nx_,
                               ny_, tref_[um], d_stage_, S(stream_));
        hip_ok(hipMemcpyAsync(T, d_stage_, N * sizeof(double), hipMemcpyDeviceToHost,
                              S(stream_)), "download");
    } else {
        const int cur = std::min(done_, n_steps(m)) & 1;
        hip_ok(hipMemcpyAsync(T, buf_[2 * um + cur], N * sizeof(double), hipMemcpyDeviceToHost,
                              S(stream_)), "download");
    }
    hip_ok(hipStreamSynchronize(S(stream_)), "download sync");
    lib_ok("download");
}

// Batches are pipelined: batch b+1 is queued (launches, ring and column
// copies into the other pinned half, an event) before the host waits for
// batch b and finalises its metrics, so the GPU does not wait for the host.
// The device ring and columns are reused in stream order.
int Ensemble::advance(int n) {
    const int start = done_;
    const int target = static_cast<int>(
        std::min<long long>(max_steps_, static_cast<long long>(done_) + std::max(n, 0)));
    int queued = done_, slot = 0;
    Pending cur{}, nxt{};
    bool have = false;
    while (have || queued < target) {
        bool have_n = false;
        if (queued < target) {
            nxt = enqueue(queued, std::min(opt_.batch, target - queued), slot);
            queued += nxt.bn;
            slot ^= 1;
            have_n = true;
        }
        if (have) complete(cur);
        cur = nxt;
        have = have_n;
    }
    return done_ - start;
}

// bn batch steps of every run and the flush; step parity is global, so the
// same sequence serves every batch start.
void Ensemble::launches(int bn) {
    const hipStream_t s = S(stream_);
    const auto* runs = static_cast<const lpbf_ensemble_run*>(d_runs_);
    for (int j = 0; j < bn; ++j)
        lpbf_launch_step_ensemble(runs, size(), nx_, ny_, opt_.rows, phys_, opt_.bc, opt_.T_bc,
                                  Tm_, d_s0_, j, s);
    lpbf_launch_step_ensemble_flush(runs, size(), nx_, ny_, opt_.rows, d_s0_, bn, s);
}

void Ensemble::capture() {
    const hipStream_t s = S(stream_);
    std::lock_guard<std::mutex> lock(setup_mutex());
    hipGraph_t g = nullptr;
    hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    launches(opt_.batch);
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        hip_ok(e, "end capture");
    }
    graph_ = g;
    lib_ok("captured launches");
    hipGraphExec_t x = nullptr;
    hip_ok(hipGraphInstantiateWithFlags(&x, g, 0), "instantiate");
    gexec_ = x;
}

// Precision 32: launches of up to `depth` steps that end at the batch end,
// at any active run's last step (then the batch is rebuilt without it) and,
// with cruise_cols, at any active run's cruise step; after a launch whose
// last step is some run's cruise step, the hot columns of every active run
// (each job copies only on its own cruise steps).  Then the records of
// the batch steps for all runs and the columns go to the pinned half `slot`.
Ensemble::Pending Ensemble::enqueue_f32(int s0, int bn, int slot) {
    const hipStream_t s = S(stream_);
    const int n = size();
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    Pending p{s0, bn, slot, bn, -1, std::vector<char>(static_cast<std::size_t>(bn), 0)};
    auto ends_at = [&](int st) {
        for (int m : active_)
            if (n_steps(m) == st + 1) return true;
        return false;
    };
    auto cruise_at = [&](int st) {
        for (int m : active_)
            if (st < n_steps(m) && cruise_[static_cast<std::size_t>(m)][static_cast<std::size_t>(st)])
                return true;
        return false;
    };
    int i = 0;
    while (i < bn && fbatch_) {
        const int limit = std::min(opt_.depth, bn - i);
        int L = 1;
        while (L < limit) {
            const int st = s0 + i + L - 1;
            if (ends_at(st) || (opt_.cruise_cols && cruise_at(st))) break;
            ++L;
        }
        if (fresh_) {   // the batch's first launch allocates its device arguments
            std::lock_guard<std::mutex> lock(setup_mutex());
            rc_ok(lpbf_f32_batch_enqueue_launch(FB(fbatch_), s0 + i, L, s), "step launch");
            fresh_ = false;
        } else {
            rc_ok(lpbf_f32_batch_enqueue_launch(FB(fbatch_), s0 + i, L, s), "step launch");
        }
        ++flaunch_;
        const int j = i + L - 1, st = s0 + j;
        p.col_at[static_cast<std::size_t>(j)] = 1;
        if (cruise_at(st)) {
            rc_ok(lpbf_launch_hot_cols_f32(static_cast<const lpbf_f32_col_job*>(d_jobs_),
                                           static_cast<int>(active_.size()), flaunch_ & 1, st, j,
                                           static_cast<long long>(n) * ny_, s),
                  "hot columns");
            p.c_lo = std::min(p.c_lo, j);
            p.c_hi = std::max(p.c_hi, j);
        }
        i += L;
        if (ends_at(st)) {
            std::lock_guard<std::mutex> lock(setup_mutex());
            rebuild_f32(st + 1);
        }
    }
    hip_ok(hipMemcpy2DAsync(h_ring_ + static_cast<std::size_t>(slot) * n * B,
                            B * sizeof(StepStats), d_stats_ + s0,
                            static_cast<std::size_t>(max_steps_) * sizeof(StepStats),
                            static_cast<std::size_t>(bn) * sizeof(StepStats),
                            static_cast<std::size_t>(n), hipMemcpyDeviceToHost, s),
           "record copy");
    const std::size_t step_cols = static_cast<std::size_t>(n) * ny_;
    double* hc = h_cols_ + slot * B * step_cols;
    if (p.c_hi >= p.c_lo)
        hip_ok(hipMemcpyAsync(hc + p.c_lo * step_cols, d_cols_ + p.c_lo * step_cols,
                              static_cast<std::size_t>(p.c_hi - p.c_lo + 1) * step_cols *
                                  sizeof(double),
                              hipMemcpyDeviceToHost, s), "column copy");
    hip_ok(hipEventRecord(static_cast<hipEvent_t>(done_ev_[slot]), s), "batch event");
    return p;
}

Ensemble::Pending Ensemble::enqueue(int s0, int bn, int slot) {
    if (opt_.precision == 32) return enqueue_f32(s0, bn, slot);
    const hipStream_t s = S(stream_);
    const int n = size();
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    Pending p{s0, bn, slot, bn, -1, {}};

    // batch steps that copy a hot column for some run: [c_lo, c_hi]
    for (int m = 0; m < n; ++m) {
        const int end = std::min(bn, n_steps(m) - s0);
        for (int i = 0; i < end; ++i)
            if (met_[static_cast<std::size_t>(m)].cruise(at(m).plan[static_cast<std::size_t>(s0 + i)])) {
                p.c_lo = std::min(p.c_lo, i);
                p.c_hi = std::max(p.c_hi, i);
            }
    }

    h_s0_[slot] = s0;
    hip_ok(hipMemcpyAsync(d_s0_, h_s0_ + slot, sizeof(int), hipMemcpyHostToDevice, s),
           "batch start");
    if (opt_.graph && bn == opt_.batch) {
        if (!gexec_) capture();
        hip_ok(hipGraphLaunch(static_cast<hipGraphExec_t>(gexec_), s), "graph launch");
        ++graph_launches_;
    } else {
        launches(bn);
    }
    lib_ok("step launches");
    const std::size_t ring_n = static_cast<std::size_t>(n) * B;
    hip_ok(hipMemcpyAsync(h_ring_ + slot * ring_n, d_ring_, ring_n * sizeof(StepStats),
                          hipMemcpyDeviceToHost, s), "ring copy");
    const std::size_t step_cols = static_cast<std::size_t>(n) * ny_;
    double* hc = h_cols_ + slot * B * step_cols;
    if (p.c_hi >= p.c_lo)
        hip_ok(hipMemcpyAsync(hc + p.c_lo * step_cols, d_cols_ + p.c_lo * step_cols,
                              static_cast<std::size_t>(p.c_hi - p.c_lo + 1) * step_cols *
                                  sizeof(double),
                              hipMemcpyDeviceToHost, s), "column copy");
    hip_ok(hipEventRecord(static_cast<hipEvent_t>(done_ev_[slot]), s), "batch event");
    return p;
}

void Ensemble::complete(Pending const& p) {
    hip_ok(hipEventSynchronize(static_cast<hipEvent_t>(done_ev_[p.slot])), "batch sync");
    lib_ok("batch");
    const int n = size();
    const std::size_t B = static_cast<std::size_t>(opt_.batch);
    const StepStats* hr = h_ring_ + p.slot * static_cast<std::size_t>(n) * B;
    const std::size_t step_cols = static_cast<std::size_t>(n) * ny_;
    const double* hc = h_cols_ + p.slot * B * step_cols;
    for (int m = 0; m < n; ++m) {
        RunMetrics& met = met_[static_cast<std::size_t>(m)];
        const int end = std::min(p.bn, n_steps(m) - p.s0);
        for (int i = 0; i < end; ++i) {
            Tick const& tk = at(m).plan[static_cast<std::size_t>(p.s0 + i)];
            const bool take = met.cruise(tk) &&
                              (opt_.precision == 64 || p.col_at[static_cast<std::size_t>(i)]);
            const double* col = take
                                    ? hc + static_cast<std::size_t>(i) * step_cols +
                                          static_cast<std::size_t>(m) * ny_
                                    : nullptr;
            const StepRecord r = met.finish(tk, hr[static_cast<std::size_t>(m) * B + i], col);
            if (opt_.keep_records) rec_[static_cast<std::size_t>(m)].push_back(r);
        }
    }
    done_ += p.bn;
}

}  // namespace lpbf

// ---- C ABI -------------------------------------------------------------------

struct lpbf_ensemble {
    lpbf::Ensemble ens;
    template <class... A>
    explicit lpbf_ensemble(A&&... a) : ens(std::forward<A>(a)...) {}
};

using lpbf::guarded;

namespace {

int bad_arg() {
    lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
    return LPBF_ERR_INVALID_ARGUMENT;
}

}  // namespace

extern "C" {

int lpbf_ensemble_create(const StepConst* k, int n_runs, int phys, double Tm, double Lx,
                         double Ly, const double* Tamb, const LaserTick* plan,
                         const double* x_csv, const unsigned char* wrapped, const int* n_steps,
                         const lpbf_run_options* opt, lpbf_ensemble** out) {
    if (!k || !out || !n_steps || n_runs < 1) return bad_arg();
    *out = nullptr;
    long long total = 0;
    for (int m = 0; m < n_runs; ++m) {
        if (n_steps[m] < 0) return bad_arg();
        total += n_steps[m];
    }
    if (total > 0 && !plan) return bad_arg();
    if (opt && opt->precision != 64 && opt->precision != 32) return bad_arg();
    return guarded([&] {
        lpbf::Domain dom;
        dom.nx = k[0].nx;
        dom.ny = k[0].ny;
        dom.Lx = Lx;
        dom.Ly = Ly;
        std::vector<lpbf::EnsembleMember> runs(static_cast<std::size_t>(n_runs));
        long long off = 0;
        for (int m = 0; m < n_runs; ++m) {
            lpbf::EnsembleMember& r = runs[static_cast<std::size_t>(m)];
            r.k = k[m];
            r.Tamb = Tamb ? Tamb[m] : 300.0;
            r.plan.resize(static_cast<std::size_t>(n_steps[m]));
            for (int s = 0; s < n_steps[m]; ++s, ++off) {
                lpbf::Tick& t = r.plan[static_cast<std::size_t>(s)];
                t.step = s;
                t.t = s * k[m].dt;
                t.x_l = plan[off].x_l;
                t.y_l = plan[off].y_l;
                t.q_scale = plan[off].q_scale;
                t.x_laser_csv = x_csv ? x_csv[off] : -1.0;
                t.wrapped = x_csv ? (wrapped ? wrapped[off] != 0 : false) : true;
                t.dump = false;
            }
        }
        lpbf::EnsembleOptions o;
        if (opt) {
            o.bc = opt->bc;
            o.T_bc = opt->T_bc;
            o.batch = opt->batch;
            o.rows = opt->rows;
            o.keep_records = opt->keep_records != 0;
            o.graph = opt->graph != 0;
            o.precision = opt->precision;
            o.depth = opt->depth;
            o.T_ref = opt->T_ref;
            o.cruise_cols = opt->cruise_cols != 0;
        }
        *out = new lpbf_ensemble(dom, phys, Tm, std::move(runs), o);
    });
}

void lpbf_ensemble_destroy(lpbf_ensemble* e) { delete e; }

int lpbf_ensemble_upload(lpbf_ensemble* e, int m, const double* T) {
    if (!e || !T || m < 0 || m >= e->ens.size()) return bad_arg();
    return guarded([&] { e->ens.upload(m, T); });
}

int lpbf_ensemble_advance(lpbf_ensemble* e, int n_steps) {
    if (!e || n_steps < 0) return -bad_arg();
    int ran = 0;
    const int rc = guarded([&] { ran = e->ens.advance(n_steps); });
    return rc ? -rc : ran;
}

int lpbf_ensemble_download(lpbf_ensemble* e, int m, double* T) {
    if (!e || !T || m < 0 || m >= e->ens.size()) return bad_arg();
    return guarded([&] { e->ens.download(m, T); });
}

int lpbf_ensemble_summary(const lpbf_ensemble* e, int m, double* out3, int* steps_done) {
    if (!e || !out3 || m < 0 || m >= e->ens.size()) return bad_arg();
    lpbf::RunMetrics const& met = e->ens.metrics(m);
    out3[0] = met.peak_T();
    out3[1] = met.peak_W();
    out3[2] = met.steady_W_sub();
    if (steps_done) *steps_done = std::min(e->ens.steps_done(), e->ens.n_steps(m));
    return 0;
}

int lpbf_ensemble_graph_launches(const lpbf_ensemble* e) {
    if (!e) return -bad_arg();
    return e->ens.graph_launches();
}

int lpbf_ensemble_records(const lpbf_ensemble* e, int m, int first, int n,
                          lpbf_step_record* out) {
    if (!e || !out || m < 0 || m >= e->ens.size() || first < 0 || n < 0 ||
        static_cast<std::size_t>(first) + static_cast<std::size_t>(n) > e->ens.records(m).size())
        return bad_arg();
    auto const& r = e->ens.records(m);
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
