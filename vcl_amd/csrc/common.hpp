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

// Sum over the wave's active lanes by DPP row operations and lane reads (no LDS traffic, unlike
// the cross-lane shuffles of wave_sum), returned to every lane.
extern "C" __device__ float __ockl_wfred_add_f32(float);
__device__ __forceinline__ float wave_sum_dpp(float v) { return __ockl_wfred_add_f32(v); }

// Inclusive prefix sum over the wave's 64 lanes (all active) by DPP: row shifts 1, 2, 4, 8 scan
// each 16-lane row, then lane 15 of rows 0 and 2 is broadcast into rows 1 and 3 and lane 31
// into rows 2 and 3.  No LDS traffic, unlike scans built on __shfl_up.
__device__ __forceinline__ u32 wave_incl_scan(u32 x) {
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

__device__ __forceinline__ u32 wave_umax(u32 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (u32)__shfl_xor((int)v, o, 64));
    return v;
}

// Fixed-point sums for LDS tables.  LDS float atomic adds run at about a tenth of the rate of
// integer ones on this GPU, so gradient sums that many lanes feed are held as 64-bit integers in
// units of 2^-S, with S chosen per workgroup from the largest |value| it adds: with
// 2^e <= max|value| < 2^(e+1), S = 51 - e puts every value times 2^S below 2^52, so up to 2^11
// adds to one sum cannot overflow, and each value is truncated by less than 2^-51 max|value|.
// The integer sums do not depend on the order of the adds.  A workgroup whose largest |value| is
// not finite (bit pattern FIXED_NONFINITE or above) keeps float atomics, so NaN and Inf propagate
// as float summation would carry them.
constexpr u32 FIXED_NONFINITE = 0x7f800000u;

__device__ __forceinline__ int fixed_scale(u32 max_abs_bits) { return 51 - ((int)(max_abs_bits >> 23) - 127); }

// y (|y| < 2^52) truncated to an integer, as a 64-bit two's-complement addend.  |y| is split
// into floor(|y| / 2^32) and the remainder |y| - 2^32 floor(|y| / 2^32), both exact in float;
// the remainder's truncation errs by less than 1.
__device__ __forceinline__ unsigned long long fixed64(float y) {
    const float a = fabsf(y);
    const float hi = floorf(a * 0x1p-32f);
    const float r = fminf(fmaf(hi, -0x1p32f, a), 0x1.fffffep31f);
    const unsigned long long q = ((unsigned long long)(u32)(int)hi << 32) | (u32)r;
    return y < 0.f ? 0ull - q : q;
}

// Open-addressing hash table in LDS: key -> NV sums (float, or 64-bit fixed point).  Used to
// pre-aggregate gradient scatter (vertex / texel grads) within a workgroup before the global
// atomics.
constexpr u32 LDS_EMPTY = 0xffffffffu;

template <int SLOTS, int NV, typename VT = float>
struct LdsTable {
    u32 key[SLOTS];
    VT val[NV][SLOTS];
};

template <int SLOTS, int NV, typename VT>
__device__ __forceinline__ void lds_table_clear(LdsTable<SLOTS, NV, VT>& t) {
    for (int i = threadIdx.x; i < SLOTS; i += blockDim.x) {
        t.key[i] = LDS_EMPTY;
#pragma unroll
        for (int k = 0; k < NV; k++) t.val[k][i] = VT(0);
    }
}

constexpr int ilog2_c(int n) { return n <= 1 ? 0 : 1 + ilog2_c(n >> 1); }

// Multiplicative (Fibonacci) hash of k into [0, 2^BITS): the top bits of k * 2^32/phi.  The low
// bits of the product depend only on the low bits of k, so keys that differ by a multiple of a
// large power of two (texel rows of a wide image) would share them.
template <int BITS>
__device__ __forceinline__ u32 fib_hash(u32 k) { return (k * 2654435761u) >> (32 - BITS); }

// Slot of key k (inserted if absent), or -1 when MAXPROBE probes find no room.  Keys only go
// from empty to their final value, so a plain read that finds k is final; the CAS is issued only
// for an empty slot (a stale "empty" is settled by the CAS).
template <int SLOTS, int NV, typename VT, int MAXPROBE = 24>
__device__ __forceinline__ int lds_table_slot(LdsTable<SLOTS, NV, VT>& t, u32 k) {
    u32 h = fib_hash<ilog2_c(SLOTS)>(k);
    for (int probe = 0; probe < MAXPROBE; probe++) {
        u32 cur = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == k) return (int)h;
        if (cur == LDS_EMPTY) {
            u32 old = atomicCAS(&t.key[h], LDS_EMPTY, k);
            if (old == LDS_EMPTY || old == k) return (int)h;
        }
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}

// Block-wide compaction step: the threads with occ append v to out[] after its first n entries,
// in thread order, and n advances by the block's count.  Every thread of the block calls it (it
// holds two barriers); wave_count is an LDS array of blockDim.x / 64 entries.
__device__ __forceinline__ void block_append(bool occ, u32 v, u32* out, u32* wave_count, u32& n) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    u64 m = __ballot(occ);
    if (lane == 0) wave_count[wave] = (u32)__popcll(m);
    __syncthreads();
    u32 off = n, total = 0;
    for (int w = 0; w < nw; w++) {
        u32 c = wave_count[w];
        off += w < wave ? c : 0u;
        total += c;
    }
    if (occ) out[off + (u32)__popcll(m & ((1ull << lane) - 1))] = v;
    n += total;
    __syncthreads();
}

// Flush of an LDS table of per-vertex clip-space grads (x, y, w) into g[4 v + {0, 1, 3}]: the
// occupied slots are compacted into used[] (SLOTS entries) and each vertex is handed to four
// consecutive lanes, so a wave's atomics cover 16 vertices' 16-byte records.  Called by every
// thread of the block after a barrier that ends the table updates.
template <int SLOTS>
__device__ __forceinline__ void lds_vertex_grad_flush(const LdsTable<SLOTS, 3>& t, u32* used, u32* wave_count, float* g) {
    u32 n = 0;
    for (int c = 0; c < SLOTS; c += blockDim.x) {
        int sl = c + threadIdx.x;
        block_append(sl < SLOTS && t.key[sl] != LDS_EMPTY, (u32)sl, used, wave_count, n);
    }
    for (u32 i = threadIdx.x; i < 4 * n; i += blockDim.x) {
        u32 sl = used[i >> 2], comp = i & 3;
        if (comp == 2) continue;
        float v = t.val[comp == 3 ? 2 : comp][sl];
        if (v != 0.f) atomicAdd(g + 4 * (size_t)t.key[sl] + comp, v);
    }
}

// The same for a fixed-point table in units of 2^-S.
template <int SLOTS>
__device__ __forceinline__ void lds_vertex_grad_flush(const LdsTable<SLOTS, 3, unsigned long long>& t, u32* used, u32* wave_count, float* g,
                                                      int S) {
    u32 n = 0;
    for (int c = 0; c < SLOTS; c += blockDim.x) {
        int sl = c + threadIdx.x;
        block_append(sl < SLOTS && t.key[sl] != LDS_EMPTY, (u32)sl, used, wave_count, n);
    }
    for (u32 i = threadIdx.x; i < 4 * n; i += blockDim.x) {
        u32 sl = used[i >> 2], comp = i & 3;
        if (comp == 2) continue;
        const long long q = (long long)t.val[comp == 3 ? 2 : comp][sl];
        if (q != 0) atomicAdd(g + 4 * (size_t)t.key[sl] + comp, ldexpf((float)q, -S));
    }
}

// The rasterizer's per-launch workspace reset as a list of ranges (tile counters, large-piece
// sub-counters and misc[] to zero, per-pixel depth keys to all ones), so that another pass can
// do it: element i of ws_clear_elem covers index i of every range; n() elements in all.
struct WsClear {
    u32* counts;
    size_t nt;
    u32* lcounts;
    size_t nl;
    u32* misc;   // 64 entries
    unsigned long long* keys;
    size_t np;
    __host__ __device__ size_t n() const {
        const size_t a = nt > nl ? nt : nl, b = np > 64 ? np : 64;
        return counts ? (a > b ? a : b) : 0;
    }
};

__device__ __forceinline__ void ws_clear_elem(const WsClear& c, size_t i) {
    if (i < c.nt) c.counts[i] = 0;
    if (i < c.nl) c.lcounts[i] = 0;
    if (i < 64) c.misc[i] = 0;
    if (i < c.np) c.keys[i] = ~0ull;
}

// Zero-fill kernel (used instead of hipMemsetAsync so sequences stay graph-safe).
__global__ void vcl_fill_u32(u32* p, size_t n, u32 v);
int vcl_fill_async(void* p, size_t bytes, hipStream_t s);
int vcl_fill_u32_async(u32* p, size_t n, u32 v, hipStream_t s);
int vcl_fill_multi_async(void* const* p, const size_t* bytes, int n, hipStream_t s);
int vcl_fill_multi_bump_async(void* const* p, const size_t* bytes, int n, double* bt, double rhom, double rhov, hipStream_t s);

// The dispatcher hands workgroup i of a launch to XCD i % 8, and each XCD has its own L2.
// xcd_swizzle(i, n, c) maps hardware block i of n to a logical block such that XCD x serves
// logical chunks x, x + 8, x + 16, ... of c consecutive blocks: with c = tiles per image row,
// each tile row sits in one L2, so the halo column a tile reads from its right neighbour (a
// full cache line for one pixel) is already there, while consecutive tile rows still spread over
// all XCDs (covered rows cost far more than background ones).  Needs n % (8 c) == 0, else the
// identity.  Any bijection gives the same results.
constexpr unsigned VCL_XCDS = 8;
__host__ __device__ __forceinline__ unsigned xcd_swizzle(unsigned i, unsigned n, unsigned c) {
    if (c == 0 || n % (VCL_XCDS * c)) return i;
    const unsigned x = i % VCL_XCDS, k = i / VCL_XCDS;
    return ((k / c) * VCL_XCDS + x) * c + k % c;
}
