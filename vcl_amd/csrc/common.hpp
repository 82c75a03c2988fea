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

// Zero-fill kernel (used instead of hipMemsetAsync so sequences stay graph-safe).
__global__ void vcl_fill_u32(u32* p, size_t n, u32 v);
int vcl_fill_async(void* p, size_t bytes, hipStream_t s);
