// SPDX-License-Identifier: 0BSD
// The GPU step task on Hedgehog itself: lpbf::hh_task::StepTileTask derives
// from hh::AbstractTask<1, HaloTile, HaloTile> as the reference's task does
// (include/step_task.hpp:37-40) and keeps its constructor (:41-51),
// execute (:53) and copy (:207-213), so reference code built on hh::Graph
// (src/main.cpp:261-266) switches by including this header in place of
// step_task.hpp and naming lpbf::hh_task::StepTileTask (with lpbf's
// Material, Process, Domain, HaloTile and StepProfile).  Each Hedgehog copy
// owns an lpbf::StepTileTask, whose step(HaloTile&) advances the tile on
// the GPU with the copy's own stream and staging buffers (created on first
// use by the worker thread) and updates the shared StepProfile.
// Requires <hedgehog/hedgehog.h>; without Hedgehog, lpbf/step_task.hpp
// with lpbf::Graph has the same calls.
#pragma once

#include <hedgehog/hedgehog.h>

#include <memory>
#include <string>

#include "step_task.hpp"

namespace lpbf {
namespace hh_task {

class StepTileTask : public hh::AbstractTask<1, HaloTile, HaloTile> {
public:
    StepTileTask(std::string const& name, std::size_t n_threads, Material mat, Process pr,
                 Domain dom, double dt, std::shared_ptr<StepProfile> prof,
                 Physics phys = Physics::Baseline, double kmult = 1.0)
        : hh::AbstractTask<1, HaloTile, HaloTile>(name, n_threads),
          mat_(mat), pr_(pr), dom_(dom), dt_(dt), prof_(prof), phys_(phys), kmult_(kmult),
          impl_(std::make_unique<lpbf::StepTileTask>(name, 1, mat, pr, dom, dt, std::move(prof),
                                                     phys, kmult)) {}

    // Replaces tile->data by the stepped tile (halo rows copied through)
    // and emits the same tile, as the reference's execute does.
    void execute(std::shared_ptr<HaloTile> tile) override {
        impl_->step(*tile);
        this->addResult(tile);
    }

    std::shared_ptr<hh::AbstractTask<1, HaloTile, HaloTile>> copy() override {
        return std::make_shared<StepTileTask>(this->name(), this->numberThreads(), mat_, pr_,
                                              dom_, dt_, prof_, phys_, kmult_);
    }

private:
    Material mat_;
    Process pr_;
    Domain dom_;
    double dt_;
    std::shared_ptr<StepProfile> prof_;
    Physics phys_;
    double kmult_;
    std::unique_ptr<lpbf::StepTileTask> impl_;
};

}  // namespace hh_task
}  // namespace lpbf
