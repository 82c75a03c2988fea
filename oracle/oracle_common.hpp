// oracle_common.hpp -- helpers shared by the oracle's translation units (test infrastructure;
// see the header of vcl_oracle.cpp).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <thread>
#include <algorithm>
#include <functional>

typedef uint32_t u32;
typedef int64_t i64;
typedef uint64_t u64;

// worker threads of parallel_for (orc_set_threads)
extern int g_threads;

// Run fn(begin, end, thread_index) over [0, n) split into contiguous chunks.
static inline void parallel_for(int n, const std::function<void(int, int, int)>& fn) {
    int nt = std::min(g_threads, std::max(n, 1));
    if (nt <= 1) { fn(0, n, 0); return; }
    std::vector<std::thread> th;
    int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &fn]() { fn(b, e, t); });
    }
    for (auto& x : th) x.join();
}

// Per-thread scatter accumulators merged in thread order (deterministic for a fixed thread
// count).  Replaces the reference's float atomics (e.g. interpolate.cu:133-135).
template <class T>
struct Scatter {
    T* dst; size_t n; int nt;
    std::vector<std::vector<T>> part;
    Scatter(T* d, size_t n_) : dst(d), n(n_), nt(std::max(1, g_threads)) {
        if (dst && nt > 1) part.assign(nt, std::vector<T>(n, T(0)));
    }
    T* buf(int t) { return (!dst) ? nullptr : (nt > 1 ? part[t].data() : dst); }
    void merge() {
        if (!dst || nt <= 1) return;
        for (int t = 0; t < nt; t++) { const T* p = part[t].data(); for (size_t i = 0; i < n; i++) dst[i] += p[i]; }
    }
};

// AddNaNcheck (common.h:181): a non-finite accumulator is reset to 0 and b dropped; a
// non-finite sum is dropped.
template <class T> static inline void add_nan_check(T& a, T b) {
    if (!std::isfinite(a)) a = T(0);
    else { T v = a + b; if (std::isfinite(v)) a = v; }
}

// max(float, double) in CUDA device code returns double (used by material.cu:166,309).
static inline double dmax(double a, double b) { return a > b ? a : b; }

