// SPDX-License-Identifier: 0BSD
// Many runs of one plate advanced together: the reference's manifest sweep
// (src/gpu/hh_gpu_main.cpp:234-290, one run_one_row per CSV row, each a
// GpuStepTask + MetricsTask pair on the same Domain) as one FP64 launch per
// step for all rows instead of one small launch per row and step.
//
// The default plate (240 x 120) is 28,800 cells: one run's step is a few
// dozen small workgroups, leaves most CUs idle and is bound by launch
// latency.  Here
// every run keeps its own fields, constants, laser plan and cruise flags in
// HBM, and batch step j of the ensemble is one lpbf_launch_step_ensemble
// over all runs (per run: the step's compute workgroups plus one that
// finishes the run's previous step into its ring).  A batch is `batch`
// such launches and a flush, replayed as one captured hipGraph, then one
// copy of all rings (and of the hot columns when a cruise step is in the
// batch) to pinned memory.  The next batch is queued before the host waits
// for this one and finalises each run's metrics in step order with the same
// RunMetrics as DeviceRun.
//
// Every cell update and hot-column copy is the one DeviceRun makes, so
// fields, the integer per-step statistics, tmax, W, W_sub and the steady
// W_sub equal a DeviceRun of the same run bit for bit; the FP64 energy sums
// merge bands of another height (chosen for the whole ensemble) and may
// differ in their last bits.  Runs may
// differ in their constants (P, V, r0, eta, dSub, kmult, dt, h), plan
// length and initial field; the grid, physics, boundary condition and Tm
// are shared.  No snapshots (the manifest loop writes none).
//
// Precision 32 (baseline physics) steps the runs still running as one
// lpbf_f32_batch: each launch advances every run by up to `depth` steps,
// ending at batch ends, at any run's last step (the batch is then rebuilt
// without the runs that ended) and, with cruise_cols, at any run's cruise
// step.  After a launch whose last step is a run's cruise step, one
// lpbf_launch_hot_cols_f32 copies the hot columns of all runs.  Fields are
// bitwise those of FP32 DeviceRuns (the FP32 field does not depend on the
// launch depth or band height), and W_sub is taken where such a DeviceRun
// ends a launch when the launch boundaries agree (equal batch, depth and
// run lengths).  Each run's theta origin is its Tamb (or T_ref if >= 0).
#pragma once

#include <vector>

#include "../lpbf_step_const.h"
#include "device_run.hpp"
#include "params.hpp"
#include "plan.hpp"

namespace lpbf {

struct EnsembleMember {
    StepConst k;              // grid equal to the ensemble's; k.kmult honoured
    std::vector<Tick> plan;   // laser and cruise gate per step (dump ignored)
    double Tamb = 300.0;      // initial uniform field
};

struct EnsembleOptions {
    int bc = LPBF_BC_MIRROR;
    double T_bc = 300.0;      // Dirichlet edge temperature
    int batch = 64;           // steps per host synchronisation
    int rows = 0;             // band height per wavefront (0: lpbf_step_ensemble_rows;
                              // precision 32: the FP32 kernel's band height, 0 = built-in)
    bool keep_records = false;
    bool graph = true;        // FP64: replay full batches as one captured hipGraph
    int precision = 64;       // 64 (FP64 ensemble launch) or 32 (batched FP32 kernel)
    int depth = 4;            // FP32: steps per launch, 1..4
    double T_ref = -1.0;      // FP32 theta origin (< 0: each run's Tamb)
    bool cruise_cols = false; // FP32: launches also end at cruise steps
};

class Ensemble {
public:
    // dom/phys/Tm as for DeviceRun.  Throws std::invalid_argument on bad
    // input and std::runtime_error on a HIP failure.
    Ensemble(Domain dom, int phys, double Tm, std::vector<EnsembleMember> runs,
             EnsembleOptions opt = {});
    ~Ensemble();
    Ensemble(Ensemble const&) = delete;
    Ensemble& operator=(Ensemble const&) = delete;

    int size() const noexcept { return static_cast<int>(runs_.size()); }
    void upload(int m, const double* T);     // run m's field (nx*ny FP64), before stepping
    // Advances the shared step counter by up to n steps (every run still
    // running takes them); returns how many.
    int advance(int n);
    int run() { return advance(max_steps_ - done_); }
    void download(int m, double* T) const;   // run m's current field

    int steps_done() const noexcept { return done_; }      // shared step counter
    int n_steps(int m) const { return static_cast<int>(at(m).plan.size()); }
    int max_steps() const noexcept { return max_steps_; }
    RunMetrics const& metrics(int m) const { return met_.at(static_cast<std::size_t>(m)); }
    std::vector<StepRecord> const& records(int m) const { return rec_.at(static_cast<std::size_t>(m)); }
    int graph_launches() const noexcept { return graph_launches_; }
    void* stream() const noexcept { return stream_; }
    int rows() const noexcept { return opt_.rows; }     // band height in use

private:
    // a queued batch: steps [s0, s0 + bn), host buffers `slot`, hot columns
    // copied for batch steps [c_lo, c_hi]
    struct Pending {
        int s0, bn, slot, c_lo, c_hi;
        std::vector<char> col_at;     // precision 32: batch step ends a launch
    };
    EnsembleMember const& at(int m) const { return runs_.at(static_cast<std::size_t>(m)); }
    Pending enqueue(int s0, int bn, int slot);
    void complete(Pending const& p);
    void launches(int bn);
    void capture();
    void release() noexcept;
    void* dalloc(std::size_t bytes, const char* what);   // device memory freed in release()
    // precision 32
    void setup_f32();
    Pending enqueue_f32(int s0, int bn, int slot);
    void rebuild_f32(int next);
    int latest_f32(int m) const;      // buffer (0/1) holding run m's latest field

    Domain dom_;
    int phys_;
    double Tm_;
    std::vector<EnsembleMember> runs_;
    EnsembleOptions opt_;
    int nx_, ny_, nblk_ = 0, max_steps_ = 0, done_ = 0, graph_launches_ = 0;
    void* stream_ = nullptr;
    std::vector<void*> dev_;          // every device allocation, freed in release()
    void* d_runs_ = nullptr;          // lpbf_ensemble_run[n] (device)
    std::vector<double*> buf_;        // 2 per run
    int* d_s0_ = nullptr;             // batch start step (device)
    int* h_s0_ = nullptr;             // pinned, two
    StepStats* d_ring_ = nullptr;     // [n][batch]
    double* d_cols_ = nullptr;        // [batch][n][ny] (step-major)
    StepStats* h_ring_ = nullptr;     // pinned, two halves
    double* h_cols_ = nullptr;        // pinned, two halves
    void* done_ev_[2] = {nullptr, nullptr};   // a half's copies have landed
    void* graph_ = nullptr;
    void* gexec_ = nullptr;
    std::vector<RunMetrics> met_;
    std::vector<std::vector<StepRecord>> rec_;
    // precision 32: theta fields (2 per run, pitch ld_), the batch of the
    // runs still running (active_), each run's latest buffer at the last
    // rebuild (fcur_) and the launches since (flaunch_), per-step records
    // [run][max_steps], the hot-column jobs of the active runs, staging for
    // FP64 uploads and downloads
    int ld_ = 0;
    std::vector<float*> fbuf_;
    std::vector<double> tref_;
    std::vector<LaserTick*> d_plan_;
    std::vector<unsigned char*> d_cruise_;
    std::vector<std::vector<unsigned char>> cruise_;
    void* fbatch_ = nullptr;          // lpbf_f32_batch*
    std::vector<int> active_;
    std::vector<int> fcur_;
    int flaunch_ = 0;
    bool fresh_ = false;              // the batch has not launched yet
    StepStats* d_stats_ = nullptr;
    void* d_jobs_ = nullptr;          // lpbf_f32_col_job[n]
    double* d_stage_ = nullptr;
};

}  // namespace lpbf
