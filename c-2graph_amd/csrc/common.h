// Shared host-side helpers: thread-local error state, a small parallel_for,
// the counter-based hash used by the generator, and timing.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <thread>
#include <vector>
#include <atomic>
#include <chrono>
#include <algorithm>
#include <climits>
#include "../../include/c2g.h"

namespace c2g {

// Thread-local message of the last failing call (c2g_last_error).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* get_error();
void clear_error();

// Returns `code` after recording a message; used as `return fail(C2G_EINVAL, ...)`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Worker count for host-side parallel loops: C2G_THREADS, else hardware
// concurrency capped at 16 (the per-GPU host share on the target nodes).
int host_threads();

// Static block partition of [0,n) over host_threads() std::threads; each
// thread gets at least `grain` items.
template <class F>
void parallel_for(int64_t n, F&& f, int nthreads = 0, int64_t grain = 4096) {
    if (nthreads <= 0) nthreads = host_threads();
    if (grain < 1) grain = 1;
    if (n <= grain || nthreads == 1) { if (n > 0) f(0, n, 0); return; }
    nthreads = (int)std::min<int64_t>(nthreads, (n + grain - 1) / grain);
    std::vector<std::thread> ts;
    ts.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        int64_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
        ts.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto& t : ts) t.join();
}

// Dynamic schedule: items [0,n) handed out `step` at a time to the threads.
template <class F>
void parallel_dyn(int64_t n, F&& f, int64_t step = 1, int nthreads = 0) {
    if (nthreads <= 0) nthreads = host_threads();
    if (n <= 0) return;
    std::atomic<int64_t> next{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t)
        ts.emplace_back([&, t] {
            for (;;) {
                int64_t b = next.fetch_add(step);
                if (b >= n) break;
                f(b, std::min(n, b + step), t);
            }
        });
    for (auto& t : ts) t.join();
}

// splitmix64 finaliser: a bijective 64-bit mixer (Steele, Lea, Flood 2014).
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// Counter-based stream: value i of stream (seed, tag) is
// hash3(seed, tag, i) = mix64(mix64(seed ^ tag*0xd1342543de82ef95) + i).
struct Stream {
    uint64_t key;
    Stream(uint64_t seed, uint64_t tag) : key(mix64(seed ^ (tag * 0xd1342543de82ef95ull))) {}
    inline uint64_t operator()(uint64_t i) const { return mix64(key + i); }
};
inline uint64_t hash3(uint64_t seed, uint64_t tag, uint64_t i) { return Stream(seed, tag)(i); }

inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace c2g
