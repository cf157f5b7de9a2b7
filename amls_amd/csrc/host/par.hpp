// Host-side fork/join helpers for the set-up work (graph, symbolic tables,
// operator layouts): plain std::thread, no pool, joined before returning.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <thread>
#include <vector>

namespace amls {

// worker threads for host set-up: the hardware threads, at most 16
inline int hostThreads() {
    const unsigned h = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(h, 16u)));
}

// fn(lo, hi, chunk) over [0, n) cut into `chunks` contiguous ranges, chunk
// ranges in order; the first exception thrown by any range is rethrown
template <typename Fn>
void parallelRanges(int64_t n, int chunks, Fn&& fn) {
    chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(chunks, n)));
    if (chunks == 1) {
        fn(int64_t(0), n, 0);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs(static_cast<size_t>(chunks));
    for (int c = 0; c < chunks; ++c) {
        const int64_t lo = n * c / chunks, hi = n * (c + 1) / chunks;
        th.emplace_back([&, lo, hi, c] {
            try {
                fn(lo, hi, c);
            } catch (...) {
                errs[static_cast<size_t>(c)] = std::current_exception();
            }
        });
    }
    for (auto& t : th) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

// fn(i) for i in [0, n) on `threads` threads taking indices in turn (for
// items of uneven cost); the first exception thrown is rethrown
template <typename Fn>
void parallelDynamic(int64_t n, int threads, Fn&& fn) {
    std::atomic<int64_t> next{0};
    parallelRanges(static_cast<int64_t>(std::max(1, threads)), std::max(1, threads),
                   [&](int64_t, int64_t, int) {
                       for (int64_t i = next++; i < n; i = next++) fn(i);
                   });
}

// runs every task of `tasks` on its own thread and joins them
template <typename Task>
void parallelTasks(std::vector<Task>& tasks) {
    parallelRanges(static_cast<int64_t>(tasks.size()), static_cast<int>(tasks.size()),
                   [&](int64_t lo, int64_t hi, int) {
                       for (int64_t i = lo; i < hi; ++i) tasks[static_cast<size_t>(i)]();
                   });
}

}  // namespace amls
