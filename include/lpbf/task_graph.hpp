// SPDX-License-Identifier: 0BSD
// A minimal single-task dataflow graph with the calling surface the
// reference drivers use from Hedgehog (hh::AbstractTask / hh::Graph as used
// in src/main.cpp:261-266, 328-348, 404-405): a task with a fixed number of
// worker threads (each running its own copy() of the task), pushData,
// getBlockingResult returning an optional variant, finishPushingData and
// waitForTermination.  Hedgehog itself is not available offline; this is
// the part of its contract the step-task interface relies on.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <variant>
#include <vector>

namespace lpbf {

template <class In, class Out>
class Graph;

template <class In, class Out>
class AbstractTask {
public:
    AbstractTask(std::string name, std::size_t n_threads)
        : name_(std::move(name)), n_threads_(n_threads < 1 ? 1 : n_threads) {}
    virtual ~AbstractTask() = default;

    virtual void execute(std::shared_ptr<In> data) = 0;
    // Needed when numberThreads() > 1: one copy per extra worker thread.
    virtual std::shared_ptr<AbstractTask> copy() { return nullptr; }
    virtual void initialize() {}
    virtual void shutdown() {}

    std::string const& name() const noexcept { return name_; }
    std::size_t numberThreads() const noexcept { return n_threads_; }

protected:
    void addResult(std::shared_ptr<Out> r) {
        if (!sink_) throw std::logic_error("addResult outside a running graph");
        sink_(std::move(r));
    }

private:
    friend class Graph<In, Out>;
    std::string name_;
    std::size_t n_threads_;
    std::function<void(std::shared_ptr<Out>)> sink_;
};

template <class In, class Out>
class Graph {
public:
    using Task = AbstractTask<In, Out>;
    using Result = std::optional<std::variant<std::shared_ptr<Out>>>;

    explicit Graph(std::string name) : name_(std::move(name)) {}
    Graph(Graph const&) = delete;
    Graph& operator=(Graph const&) = delete;
    ~Graph() {
        finishPushingData();
        join();
    }

    void inputs(std::shared_ptr<Task> t) { set_task(std::move(t)); }
    void outputs(std::shared_ptr<Task> t) { set_task(std::move(t)); }

    void executeGraph() {
        if (!task_) throw std::logic_error("graph has no task");
        if (!workers_.empty()) return;
        std::vector<std::shared_ptr<Task>> copies{task_};
        for (std::size_t i = 1; i < task_->numberThreads(); ++i) {
            auto c = task_->copy();
            if (!c) throw std::logic_error("task " + task_->name() + " needs copy() for >1 thread");
            copies.push_back(std::move(c));
        }
        running_ = copies.size();
        for (auto& c : copies) {
            c->sink_ = [this](std::shared_ptr<Out> r) { push_result(std::move(r)); };
            workers_.emplace_back([this, c] { work(c); });
        }
    }

    void pushData(std::shared_ptr<In> d) {
        {
            std::lock_guard<std::mutex> g(m_);
            if (closed_) throw std::logic_error("pushData after finishPushingData");
            in_.push_back(std::move(d));
        }
        cv_in_.notify_one();
    }

    void finishPushingData() {
        {
            std::lock_guard<std::mutex> g(m_);
            closed_ = true;
        }
        cv_in_.notify_all();
    }

    // Next result, or nullopt once input is closed and every result is taken.
    Result getBlockingResult() {
        std::unique_lock<std::mutex> g(m_);
        cv_out_.wait(g, [this] { return !out_.empty() || error_ || (closed_ && running_ == 0); });
        if (error_) std::rethrow_exception(error_);
        if (out_.empty()) return std::nullopt;
        auto r = std::move(out_.front());
        out_.pop_front();
        return Result{std::in_place, std::move(r)};
    }

    void waitForTermination() {
        join();
        std::lock_guard<std::mutex> g(m_);
        if (error_) std::rethrow_exception(error_);
    }

    std::string const& name() const noexcept { return name_; }

private:
    void set_task(std::shared_ptr<Task> t) {
        if (task_ && task_ != t) throw std::logic_error("this graph holds a single task");
        task_ = std::move(t);
    }

    void push_result(std::shared_ptr<Out> r) {
        {
            std::lock_guard<std::mutex> g(m_);
            out_.push_back(std::move(r));
        }
        cv_out_.notify_one();
    }

    void work(std::shared_ptr<Task> t) {
        try {
            t->initialize();
            for (;;) {
                std::shared_ptr<In> d;
                {
                    std::unique_lock<std::mutex> g(m_);
                    cv_in_.wait(g, [this] { return !in_.empty() || closed_ || error_; });
                    if (error_ || in_.empty()) break;
                    d = std::move(in_.front());
                    in_.pop_front();
                }
                t->execute(std::move(d));
            }
            t->shutdown();
        } catch (...) {
            std::lock_guard<std::mutex> g(m_);
            if (!error_) error_ = std::current_exception();
        }
        {
            std::lock_guard<std::mutex> g(m_);
            --running_;
        }
        cv_out_.notify_all();
        cv_in_.notify_all();
    }

    void join() {
        for (auto& w : workers_)
            if (w.joinable()) w.join();
    }

    std::string name_;
    std::shared_ptr<Task> task_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_in_, cv_out_;
    std::deque<std::shared_ptr<In>> in_;
    std::deque<std::shared_ptr<Out>> out_;
    bool closed_ = false;
    std::size_t running_ = 0;
    std::exception_ptr error_;
};

}  // namespace lpbf
