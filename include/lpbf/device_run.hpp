// SPDX-License-Identifier: 0BSD
// Device-resident run orchestration: the reference GpuStepTask batched path
// (include/gpu_step_task.hpp:135-239) and MetricsTask finalisation
// (:375-411), without Hedgehog.
//
// The field stays in HBM (double buffer).  Each batch of up to `batch` steps
// is launched back to back on one stream: the step with fused statistics,
// the per-step merge into a device ring, and on cruise steps the hot-column
// copy; then one device-to-host copy of the rings and one stream
// synchronisation, after which the per-step metrics (W, W_sub, running
// maxima, the cruise-gated steady W_sub) are finalised in step order.
//
// Precision 64 runs the FP64 parity kernels (one step per launch);
// precision 32 runs the FP32 path with `depth` steps per launch, so the hot
// column, and with it W_sub, exists only for the last step of each launch.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../lpbf_step_const.h"
#include "params.hpp"
#include "plan.hpp"

namespace lpbf {

struct RunOptions {
    int precision = 64;            // 64 or 32
    int bc = LPBF_BC_MIRROR;
    double T_bc = 300.0;
    Physics phys = Physics::Baseline;   // extended physics: precision 64 only
    double kmult = 1.0;
    int batch = 64;                // steps per host synchronisation
    int depth = 4;                 // FP32 steps per launch (1..4)
    int rows = 0;                  // FP32 band height (0 = built-in)
    double T_ref = -1.0;           // FP32 theta origin (< 0: Tamb)
};

struct StepRecord {
    Tick tick;
    StepStats stats;
    double W_um = 0.0;
    double W_sub_um = 0.0;
    bool has_col = false;
};

class DeviceRun {
public:
    DeviceRun(Material mat, Process pr, Domain dom, double dt, std::vector<Tick> plan,
              RunOptions opt);
    ~DeviceRun();
    DeviceRun(DeviceRun const&) = delete;
    DeviceRun& operator=(DeviceRun const&) = delete;

    void upload(const double* T);          // nx*ny row-major FP64
    void noise(std::uint64_t seed, double base, double amp);
    int advance(int n_steps);              // returns steps actually run
    void download(double* T) const;        // current field as FP64 T

    int steps_done() const noexcept { return done_; }
    double peak_T() const noexcept { return peak_T_; }
    double peak_W() const noexcept { return peak_W_; }
    double steady_W_sub() const noexcept { return steady_; }
    std::vector<StepRecord> const& records() const noexcept { return rec_; }
    void* stream() const noexcept { return stream_; }

private:
    void batch(int bn);

    Material mat_;
    Process pr_;
    Domain dom_;
    double dt_;
    std::vector<Tick> plan_;
    RunOptions opt_;
    StepConst k_{};
    int nx_, ny_, ld_ = 0, nblk_ = 0;
    void* stream_ = nullptr;
    void* d_a_ = nullptr;          // current field
    void* d_b_ = nullptr;
    void* d_plan_ = nullptr;       // LaserTick[n_steps]
    void* d_blk_ = nullptr;        // block records / FP32 workspace
    StepStats* d_ring_ = nullptr;  // per-step stats of the batch
    double* d_cols_ = nullptr;     // per-step hot columns of the batch
    std::vector<StepStats> h_ring_;
    std::vector<double> h_cols_;
    std::vector<StepRecord> rec_;
    int done_ = 0;
    double peak_T_, peak_W_ = 0.0, steady_ = 0.0;
};

}  // namespace lpbf
