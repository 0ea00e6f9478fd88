// Small host parallel loop: [0, n) split into one contiguous range per
// thread; the first exception thrown by a worker is rethrown on the caller.
#pragma once
#include <algorithm>
#include <cstdint>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace igr {

template <class F>
void parallel_ranges(int64_t n, int threads, F&& fn) {
    threads = std::max(1, threads);
    if (n <= 0) return;
    if (threads == 1 || n < 2) {
        fn((int64_t)0, n, 0);
        return;
    }
    std::exception_ptr err;
    std::mutex mu;
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        pool.emplace_back([&, a, b, t]() {
            try {
                fn(a, b, t);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) err = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

} // namespace igr
