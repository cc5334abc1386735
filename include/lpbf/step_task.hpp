// SPDX-License-Identifier: 0BSD
// The step-task drop-in: StepTileTask advances one HaloTile by one explicit
// Euler step, on the GPU.
//
// Same constructor, execute() contract and copy() semantics as the reference
// include/step_task.hpp:37-224: execute reads the tile including its halo
// rows, replaces tile->data by the stepped tile (halo rows copied through),
// emits the same shared_ptr, and touches nothing else; the laser comes from
// tile->x_laser / y_laser / q_scale and the global row of the first owned row
// from tile->y0.  Each worker copy owns a HIP stream and device staging
// buffers; the update itself is lpbf_launch_step_tile (FP64, the parity
// kernel).  StepProfile keeps the reference counters (relaxed atomics).
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>

#include "params.hpp"
#include "task_graph.hpp"
#include "tile.hpp"

namespace lpbf {

struct StepProfile {
    std::atomic<std::int64_t> compute_ns{0};   // wall time inside execute()
    std::atomic<std::int64_t> alloc_ns{0};     // host result buffer allocation
    std::atomic<std::int64_t> stencil_ns{0};   // upload + kernel + download
    std::atomic<std::int64_t> n_calls{0};
};

class StepTileTask : public AbstractTask<HaloTile, HaloTile> {
public:
    StepTileTask(std::string const& name, std::size_t n_threads, Material mat, Process pr,
                 Domain dom, double dt, std::shared_ptr<StepProfile> prof,
                 Physics phys = Physics::Baseline, double kmult = 1.0) noexcept;
    ~StepTileTask() override;

    void execute(std::shared_ptr<HaloTile> tile) override;
    // The update of execute() without emitting: replaces tile.data by the
    // stepped tile.  For adapters onto other task runtimes.
    void step(HaloTile& tile);
    std::shared_ptr<AbstractTask<HaloTile, HaloTile>> copy() override;
    void shutdown() override;

private:
    void reserve(std::size_t n_values);
    void release() noexcept;

    Material mat_;
    Process pr_;
    Domain dom_;
    double dt_;
    std::shared_ptr<StepProfile> prof_;
    Physics phys_;
    double kmult_;
    StepConst k_;
    // per-copy device state, created on first use by the worker thread
    void* stream_ = nullptr;
    double* d_in_ = nullptr;
    double* d_out_ = nullptr;
    std::size_t cap_ = 0;
};

}  // namespace lpbf
