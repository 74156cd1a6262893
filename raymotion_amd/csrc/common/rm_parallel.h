// Chunked parallel loop for the host-side O(triangles) passes.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace rmpar {

// Threads used by parallel_chunks: the hardware threads, at most 16.
inline unsigned pool_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, h ? h : 1u));
}

// Runs fn(c) for every chunk c in [0, n); chunks are claimed dynamically, so
// fn must only write data owned by its chunk.
template <class Fn>
void parallel_chunks(size_t n, Fn fn) {
    const size_t T = std::min<size_t>(pool_threads(), n);
    if (T <= 1) {
        for (size_t c = 0; c < n; ++c) fn(c);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    pool.reserve(T);
    for (size_t t = 0; t < T; ++t)
        pool.emplace_back([&] {
            for (size_t c; (c = next.fetch_add(1)) < n;) fn(c);
        });
    for (auto& th : pool) th.join();
}

}  // namespace rmpar
