// SPDX-License-Identifier: 0BSD
// StepTileTask: the reference step-task interface (include/step_task.hpp:
// 37-224) backed by the gfx950 tile kernel.
#include "lpbf/step_task.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpbf_kernels.h"

namespace lpbf {

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("StepTileTask: ") + what + ": " + hipGetErrorString(e));
}

std::int64_t ns_between(std::chrono::steady_clock::time_point a,
                        std::chrono::steady_clock::time_point b) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
}

}  // namespace

StepTileTask::StepTileTask(std::string const& name, std::size_t n_threads, Material mat,
                           Process pr, Domain dom, double dt, std::shared_ptr<StepProfile> prof,
                           Physics phys, double kmult) noexcept
    : AbstractTask<HaloTile, HaloTile>(name, n_threads),
      mat_(mat), pr_(pr), dom_(dom), dt_(dt), prof_(std::move(prof)), phys_(phys),
      kmult_(kmult) {
    k_ = make_const(mat_, pr_, dom_, dt_);
    k_.kmult = kmult_;
}

StepTileTask::~StepTileTask() { release(); }

void StepTileTask::release() noexcept {
    if (d_in_) hipFree(d_in_);
    if (d_out_) hipFree(d_out_);
    if (stream_) hipStreamDestroy(static_cast<hipStream_t>(stream_));
    d_in_ = d_out_ = nullptr;
    stream_ = nullptr;
    cap_ = 0;
}

void StepTileTask::shutdown() { release(); }

void StepTileTask::reserve(std::size_t n) {
    if (!stream_) {
        hipStream_t s;
        hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        stream_ = s;
    }
    if (n <= cap_) return;
    if (d_in_) hip_ok(hipFree(d_in_), "free");
    if (d_out_) hip_ok(hipFree(d_out_), "free");
    d_in_ = d_out_ = nullptr;
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_in_), n * sizeof(double)), "malloc");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_out_), n * sizeof(double)), "malloc");
    cap_ = n;
}

void StepTileTask::execute(std::shared_ptr<HaloTile> tile) {
    step(*tile);
    this->addResult(tile);
}

void StepTileTask::step(HaloTile& tile_ref) {
    HaloTile* tile = &tile_ref;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const int nx = tile->nx;
    const int rows = tile->total_rows();
    const std::size_t n = static_cast<std::size_t>(rows) * nx;
    if (nx != dom_.nx || tile->owned_rows() < 1 || tile->data.size() != n)
        throw std::invalid_argument("StepTileTask: tile shape does not match the domain");

    const auto ta0 = clk::now();
    std::vector<double> T_new(n);
    const auto ta1 = clk::now();

    reserve(n);
    hipStream_t s = static_cast<hipStream_t>(stream_);
    StepConst k = k_;
    k.x_l = tile->x_laser;
    k.y_l = tile->y_laser;
    k.q_scale = tile->q_scale;
    hip_ok(hipMemcpyAsync(d_in_, tile->data.data(), n * sizeof(double), hipMemcpyHostToDevice, s),
           "upload");
    lpbf_launch_step_tile(d_in_, d_out_, tile->owned_rows(), tile->y0, &k,
                          phys_ == Physics::Extended ? LPBF_PHYS_EXTENDED : LPBF_PHYS_BASELINE, s);
    const int err = lpbf_get_last_error();
    if (err) throw std::runtime_error(std::string("StepTileTask: ") + lpbf_error_string(err));
    hip_ok(hipMemcpyAsync(T_new.data(), d_out_, n * sizeof(double), hipMemcpyDeviceToHost, s),
           "download");
    hip_ok(hipStreamSynchronize(s), "sync");
    const auto ts1 = clk::now();

    tile->data = std::move(T_new);

    const auto t1 = clk::now();
    if (prof_) {
        prof_->alloc_ns.fetch_add(ns_between(ta0, ta1), std::memory_order_relaxed);
        prof_->stencil_ns.fetch_add(ns_between(ta1, ts1), std::memory_order_relaxed);
        prof_->compute_ns.fetch_add(ns_between(t0, t1), std::memory_order_relaxed);
        prof_->n_calls.fetch_add(1, std::memory_order_relaxed);
    }
}

std::shared_ptr<AbstractTask<HaloTile, HaloTile>> StepTileTask::copy() {
    return std::make_shared<StepTileTask>(this->name(), this->numberThreads(), mat_, pr_, dom_, dt_,
                                          prof_, phys_, kmult_);
}

}  // namespace lpbf
