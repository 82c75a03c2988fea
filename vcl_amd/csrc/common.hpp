// common.hpp -- shared device helpers for the vcl_amd HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include "../../include/vcl.h"

typedef uint32_t u32;
typedef int64_t i64;
typedef uint64_t u64;

#define VCL_HIP_CHECK(call)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return VCL_ERR_HIP;        \
    } while (0)

#define VCL_REQUIRE(cond)                                \
    do {                                                 \
        if (!(cond)) return VCL_ERR_INVALID;             \
    } while (0)

#define VCL_LAUNCH_CHECK() VCL_HIP_CHECK(hipGetLastError())

static inline hipStream_t as_hip(vcl_stream_t s) { return (hipStream_t)s; }
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// AddNaNcheck (common.h:181): non-finite accumulator -> 0 (b dropped); non-finite sum dropped.
__device__ __forceinline__ void add_nan_check(float& a, float b) {
    if (!isfinite(a)) a = 0.f;
    else { float v = a + b; if (isfinite(v)) a = v; }
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// Wave64 sum via DPP/permute (all lanes end with the total).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Open-addressing hash table in LDS: key -> NV float sums (ds_add_f32).  Used to pre-aggregate
// gradient scatter (vertex / texel grads) within a workgroup before the global atomics.
constexpr u32 LDS_EMPTY = 0xffffffffu;

template <int SLOTS, int NV>
struct LdsTable {
    u32 key[SLOTS];
    float val[NV][SLOTS];
};

template <int SLOTS, int NV>
__device__ __forceinline__ void lds_table_clear(LdsTable<SLOTS, NV>& t) {
    for (int i = threadIdx.x; i < SLOTS; i += blockDim.x) {
        t.key[i] = LDS_EMPTY;
#pragma unroll
        for (int k = 0; k < NV; k++) t.val[k][i] = 0.f;
    }
}

constexpr int ilog2_c(int n) { return n <= 1 ? 0 : 1 + ilog2_c(n >> 1); }

// Multiplicative (Fibonacci) hash of k into [0, 2^BITS): the top bits of k * 2^32/phi.  The low
// bits of the product depend only on the low bits of k, so keys that differ by a multiple of a
// large power of two (texel rows of a wide image) would share them.
template <int BITS>
__device__ __forceinline__ u32 fib_hash(u32 k) { return (k * 2654435761u) >> (32 - BITS); }

// Slot of key k (inserted if absent), or -1 when MAXPROBE probes find no room.
template <int SLOTS, int NV, int MAXPROBE = 24>
__device__ __forceinline__ int lds_table_slot(LdsTable<SLOTS, NV>& t, u32 k) {
    u32 h = fib_hash<ilog2_c(SLOTS)>(k);
    for (int probe = 0; probe < MAXPROBE; probe++) {
        u32 old = atomicCAS(&t.key[h], LDS_EMPTY, k);
        if (old == LDS_EMPTY || old == k) return (int)h;
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}

// Zero-fill kernel (used instead of hipMemsetAsync so sequences stay graph-safe).
__global__ void vcl_fill_u32(u32* p, size_t n, u32 v);
int vcl_fill_async(void* p, size_t bytes, hipStream_t s);
int vcl_fill_u32_async(u32* p, size_t n, u32 v, hipStream_t s);
int vcl_fill_multi_async(void* const* p, const size_t* bytes, int n, hipStream_t s);
