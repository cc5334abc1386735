// SPDX-License-Identifier: 0BSD
// Device-resident run orchestration: what the reference GpuStepTask batched
// path (include/gpu_step_task.hpp:135-239, mode 2, fused and pinned) and
// MetricsTask (:375-411) do together, without Hedgehog.
//
// The field stays in HBM (double buffer).  A batch of up to `batch` steps is
// launched back to back on the run's own stream: the step with fused
// statistics, the per-step merge into a device ring, on cruise steps the
// hot-column copy into a column ring, and on dump steps an asynchronous copy
// of the field to host memory; then one device-to-host copy of the rings
// into pinned staging.  The next batch is queued before the host waits for
// that copy and finalises the per-step metrics (W, W_sub, running maxima,
// the cruise-gated steady W_sub) in step order, exactly as MetricsTask
// does, so the GPU does not wait for the host.
//
// Precision 64 runs the FP64 parity kernels, one step per launch, so every
// per-step value equals the reference's.  A full batch that starts at a
// multiple of `batch` (even) and holds no snapshot is launched as one
// captured hipGraph: the laser tick and cruise gate of each step come from
// batch copies in device memory, refreshed before each replay, so the same
// graph replays for every such batch (kernel nodes only).  With f64_depth > 1
// (large plates) the FP64 steps run on the temporally blocked FP64 kernel
// instead, up to f64_depth steps per launch, a launch ending at every cruise
// or dump step (the hot column and the snapshot need that step's field); the
// fields are bitwise those of single steps.  Precision 32 runs
// the temporally blocked FP32 kernel with up to `depth` steps per launch
// (the extended physics: lpbf_launch_step_f32_ext, one step per launch); a launch never
// crosses a dump step or a batch end, and the hot column (hence W_sub) is
// taken at the last step of each launch only, or, with cruise_cols, at every
// cruise step (launches also end there).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "../lpbf_step_const.h"
#include "params.hpp"
#include "plan.hpp"

namespace lpbf {

struct RunOptions {
    int precision = 64;            // 64 or 32
    int bc = LPBF_BC_MIRROR;
    double T_bc = 300.0;           // Dirichlet edge temperature
    int batch = 64;                // steps per host synchronisation
    int depth = 4;                 // FP32 steps per launch (1..4)
    int rows = 0;                  // band height of the blocked kernels (0 = built-in)
    double T_ref = -1.0;           // FP32 theta origin (< 0: Tamb)
    bool keep_records = true;      // keep one StepRecord per step
    bool graph = true;             // FP64: replay full batches as one captured hipGraph
    int f64_depth = 1;             // FP64 steps per launch (1..4); above 1, the
                                   // temporally blocked kernel, graph off
    bool cruise_cols = false;      // FP32: launches also end at cruise steps, so
                                   // every cruise step gets its hot column
};

struct StepRecord {
    int step = 0;
    bool has_col = false;          // W_sub was computed from a hot column
    StepStats stats{};
    double W_um = 0.0;
    double W_sub_um = 0.0;
};

// Per-step metrics of one run, finalised on the host in step order from the
// step's merged statistics and, on cruise steps, the column through the
// hottest cell (MetricsTask::execute, gpu_step_task.hpp:375-404, same
// expressions): the molten-row width W, the sub-grid width W_sub, running
// maxima, and the steady W_sub (the maximum over cruise steps).
class RunMetrics {
public:
    RunMetrics(Domain dom, double Tm) : dom_(dom), Tm_(Tm) {}
    // cruise gate: first pass, laser between 45 % and 85 % of the plate
    bool cruise(Tick const& t) const noexcept {
        return !t.wrapped && t.x_laser_csv >= 0.45 * dom_.Lx && t.x_laser_csv <= 0.85 * dom_.Lx;
    }
    // `col` is the ny temperatures of the hot column, or null (no column)
    StepRecord finish(Tick const& t, StepStats const& st, const double* col);
    double peak_T() const noexcept { return peak_T_; }
    double peak_W() const noexcept { return peak_W_; }
    double steady_W_sub() const noexcept { return steady_; }

private:
    Domain dom_;
    double Tm_;
    double peak_T_ = 0.0, peak_W_ = 0.0, steady_ = 0.0;
};

class DeviceRun {
public:
    // dom/k/phys/Tm/plan/Tamb have the meaning of the reference GpuStepTask
    // constructor arguments (gpu_step_task.hpp:63-66); k->kmult is honoured.
    // Throws std::invalid_argument / std::runtime_error on bad input or a
    // HIP failure.
    DeviceRun(Domain dom, StepConst k, int phys, double Tm, std::vector<Tick> plan, double Tamb,
              RunOptions opt = {});
    ~DeviceRun();
    DeviceRun(DeviceRun const&) = delete;
    DeviceRun& operator=(DeviceRun const&) = delete;

    // Called for every dump step (plan[s].dump) in step order, after the batch
    // that contains it; `field` is the FP64 row-major T after step s.
    void on_snapshot(std::function<void(int step, std::vector<double>&& field)> cb) {
        snap_cb_ = std::move(cb);
    }

    void upload(const double* T);                                 // nx*ny FP64
    void noise(std::uint64_t seed, double base, double amp);      // T = base + U(-amp, amp)
    int advance(int n_steps);      // runs up to n_steps more steps; returns how many
    int run() { return advance(static_cast<int>(plan_.size()) - done_); }
    void download(double* T) const;                               // current T, FP64

    int steps_done() const noexcept { return done_; }
    int n_steps() const noexcept { return static_cast<int>(plan_.size()); }
    double peak_T() const noexcept { return met_.peak_T(); }
    double peak_W() const noexcept { return met_.peak_W(); }
    double steady_W_sub() const noexcept { return met_.steady_W_sub(); }
    std::vector<StepRecord> const& records() const noexcept { return rec_; }
    StepRecord const& last() const noexcept { return last_; }
    void* stream() const noexcept { return stream_; }
    int graph_launches() const noexcept { return graph_launches_; }

private:
    struct Snap {
        int step;
        std::vector<double> field;
    };
    // a queued batch: steps [s0, s0 + bn), pinned half `slot`
    struct Pending {
        int s0 = 0, bn = 0, slot = 0;
        std::vector<char> has_col;   // batch step copied a hot column
        std::vector<Snap> snaps;     // snapshot copies in flight
    };
    Pending enqueue(int s0, int bn, int slot);
    void complete(Pending& p);
    bool graph_batch(int s0, int bn);
    void capture(int parity);
    void release() noexcept;
    void release_locked() noexcept;
    bool cruise(Tick const& t) const noexcept { return met_.cruise(t); }
    void finish_step(Tick const& t, StepStats const& st, const double* col);

    Domain dom_;
    StepConst k_;
    int phys_;
    double Tm_;
    std::vector<Tick> plan_;
    double Tamb_;
    RunOptions opt_;
    std::function<void(int, std::vector<double>&&)> snap_cb_;

    int nx_, ny_, ld_ = 0, nblk_ = 0;
    double T_ref_ = 0.0;
    void* stream_ = nullptr;
    void* d_a_ = nullptr;          // current field (FP64 T or FP32 theta)
    void* d_b_ = nullptr;
    LaserTick* d_plan_ = nullptr;  // FP32 path: laser per step
    void* d_blk_ = nullptr;        // FP64 block records / blocked-kernel workspace
    StepStats* d_ring_ = nullptr;  // per-step stats of a batch
    double* d_cols_ = nullptr;     // per-step hot columns of a batch
    double* d_snap_ = nullptr;     // FP32 path: FP64 staging for snapshots
    StepStats* h_ring_ = nullptr;  // pinned, two halves
    double* h_cols_ = nullptr;     // pinned, two halves
    void* done_ev_[2] = {nullptr, nullptr};   // a half's copies have landed
    // FP64 graph replay: laser ticks and cruise flags of the whole run, the
    // current batch's copies the captured launches read, one captured batch
    // per physical buffer holding the current field
    unsigned char* d_cruise_ = nullptr;
    LaserTick* d_bticks_ = nullptr;
    unsigned char* d_bcruise_ = nullptr;
    void* buf0_ = nullptr;
    void* graph_[2] = {nullptr, nullptr};
    void* gexec_[2] = {nullptr, nullptr};
    int graph_launches_ = 0;
    std::vector<StepRecord> rec_;
    StepRecord last_;
    int done_ = 0;
    RunMetrics met_;
};

}  // namespace lpbf
