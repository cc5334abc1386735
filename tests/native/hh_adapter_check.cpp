// Compiles include/lpbf/hh_step_task.hpp against the Hedgehog task
// interface (tests/native/hh_standin) and checks, on the CPU, what needs no
// GPU: the adapter is an hh::AbstractTask<1, HaloTile, HaloTile> built with
// the reference's constructor arguments, and copy() gives an independent task
// with the same name, thread count and shared profile.
#include <cstdio>
#include <memory>
#include <type_traits>

#include "lpbf/hh_step_task.hpp"

int main() {
    using Task = lpbf::hh_task::StepTileTask;
    using Base = hh::AbstractTask<1, lpbf::HaloTile, lpbf::HaloTile>;
    static_assert(std::is_base_of<Base, Task>::value, "adapter is a Hedgehog task");
    lpbf::Material mat;
    lpbf::Process pr;
    lpbf::Domain dom;
    auto prof = std::make_shared<lpbf::StepProfile>();
    const double dt = lpbf::dt_cfl(mat, dom);
    auto t = std::make_shared<Task>("Step", 4, mat, pr, dom, dt, prof, lpbf::Physics::Extended,
                                    2.5);
    std::shared_ptr<Base> c = t->copy();
    if (!c || c.get() == t.get()) { std::printf("FAIL copy\n"); return 1; }
    if (c->name() != "Step" || c->numberThreads() != 4) { std::printf("FAIL fields\n"); return 1; }
    if (prof.use_count() < 3) { std::printf("FAIL profile not shared\n"); return 1; }
    std::printf("OK\n");
    return 0;
}
