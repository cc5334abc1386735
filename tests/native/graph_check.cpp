// Exercises lpbf::Graph / AbstractTask (include/lpbf/task_graph.hpp) on the
// CPU: several worker copies, every pushed item comes back exactly once,
// getBlockingResult ends with nullopt after finishPushingData, and an
// exception thrown in execute() reaches the caller.
#include <atomic>
#include <cstdio>
#include <memory>
#include <set>
#include <stdexcept>

#include "lpbf/task_graph.hpp"

struct Item { int v; };

class Square : public lpbf::AbstractTask<Item, Item> {
public:
    Square(std::size_t n, std::shared_ptr<std::atomic<int>> copies, int fail_at)
        : AbstractTask("Square", n), copies_(copies), fail_at_(fail_at) {}
    void execute(std::shared_ptr<Item> d) override {
        if (d->v == fail_at_) throw std::runtime_error("planted");
        d->v = d->v * d->v;
        addResult(d);
    }
    std::shared_ptr<AbstractTask<Item, Item>> copy() override {
        copies_->fetch_add(1);
        return std::make_shared<Square>(numberThreads(), copies_, fail_at_);
    }
private:
    std::shared_ptr<std::atomic<int>> copies_;
    int fail_at_;
};

int main() {
    auto copies = std::make_shared<std::atomic<int>>(0);
    {
        lpbf::Graph<Item, Item> g("g");
        auto t = std::make_shared<Square>(4, copies, -1);
        g.inputs(t); g.outputs(t); g.executeGraph();
        for (int i = 0; i < 1000; ++i) g.pushData(std::make_shared<Item>(Item{i}));
        g.finishPushingData();
        std::multiset<int> got;
        while (auto r = g.getBlockingResult()) got.insert(std::get<std::shared_ptr<Item>>(*r)->v);
        g.waitForTermination();
        if (got.size() != 1000) { std::printf("FAIL count %zu\n", got.size()); return 1; }
        for (int i = 0; i < 1000; ++i)
            if (got.count(i * i) != 1) { std::printf("FAIL missing %d\n", i); return 1; }
        if (copies->load() != 3) { std::printf("FAIL copies %d\n", copies->load()); return 1; }
    }
    {
        lpbf::Graph<Item, Item> g("g");
        auto t = std::make_shared<Square>(2, copies, 7);
        g.inputs(t); g.outputs(t); g.executeGraph();
        for (int i = 0; i < 20; ++i) g.pushData(std::make_shared<Item>(Item{i}));
        g.finishPushingData();
        bool caught = false;
        try { while (g.getBlockingResult()) {} } catch (std::runtime_error const&) { caught = true; }
        try { g.waitForTermination(); } catch (std::runtime_error const&) { caught = true; }
        if (!caught) { std::printf("FAIL no exception\n"); return 1; }
    }
    std::printf("OK\n");
    return 0;
}
