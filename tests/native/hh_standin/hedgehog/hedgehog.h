// Test stand-in for the part of Hedgehog's task API that the reference's
// StepTileTask uses (include/step_task.hpp:37-51, 207-213 of the reference):
// hh::AbstractTask<1, In, Out> with a (name, number of threads) constructor,
// a virtual execute(shared_ptr<In>), addResult, copy, name() and
// numberThreads().  It lets tests/native/hh_adapter_check.cpp compile
// include/lpbf/hh_step_task.hpp against that interface; it is not Hedgehog.
#pragma once

#include <cstddef>
#include <functional>
#include <memory>
#include <string>
#include <utility>

namespace hh {

template <std::size_t Separator, class... Types>
class AbstractTask;

template <class In, class Out>
class AbstractTask<1, In, Out> {
public:
    AbstractTask(std::string const& name, std::size_t n_threads)
        : name_(name), n_threads_(n_threads) {}
    virtual ~AbstractTask() = default;
    virtual void execute(std::shared_ptr<In>) = 0;
    virtual std::shared_ptr<AbstractTask<1, In, Out>> copy() { return nullptr; }
    std::string const& name() const { return name_; }
    std::size_t numberThreads() const { return n_threads_; }
    void addResult(std::shared_ptr<Out> r) {
        if (sink) sink(std::move(r));
    }
    std::function<void(std::shared_ptr<Out>)> sink;   // where results go (test harness)

private:
    std::string name_;
    std::size_t n_threads_;
};

}  // namespace hh
