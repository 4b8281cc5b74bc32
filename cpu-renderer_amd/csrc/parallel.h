// parallel.h — fork/join over contiguous chunks for the host-side scene work
// (BVH build, flattening, scene cache files).
#pragma once
#include <sched.h>

#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace cr {

// Host threads this process may run on (its CPU affinity, which can be far
// below the machine's CPU count in a container).
inline size_t host_threads() {
    static const size_t n = [] {
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) return size_t(std::max(1, CPU_COUNT(&set)));
        return size_t(std::max(1u, std::thread::hardware_concurrency()));
    }();
    return n;
}

// Chunks parallel_chunks(n, min_chunk, f) splits [0, n) into.
inline size_t chunk_count(size_t n, size_t min_chunk) {
    return std::max<size_t>(1, std::min(host_threads(), n / std::max<size_t>(min_chunk, 1)));
}

// Runs f(begin, end, chunk) over [0, n) in up to host_threads() contiguous
// chunks of at least min_chunk items (one chunk: on this thread).
template <class F>
void parallel_chunks(size_t n, size_t min_chunk, F&& f) {
    const size_t chunks = chunk_count(n, min_chunk);
    if (chunks == 1) {
        f(size_t(0), n, size_t(0));
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(chunks - 1);
    for (size_t c = 1; c < chunks; ++c) pool.emplace_back([&, c] { f(c * n / chunks, (c + 1) * n / chunks, c); });
    f(size_t(0), n / chunks, size_t(0));
    for (auto& t : pool) t.join();
}

}  // namespace cr
