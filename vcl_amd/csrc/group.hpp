// group.hpp -- shared launches over several fit requests.
//
// A kernel of the fit step is written as a functor whose body sees its workgroup index and grid
// size through a VB value instead of blockIdx / gridDim.  The plain kernel calls the functor
// with the hardware values.  vcl_group_kernel<F, P...> runs the same body for several requests
// in one launch: the launch's workgroups are the concatenation of each request's own grid, and
// each workgroup finds its request in a device-side work table (the prefix of the requests'
// workgroup counts), takes that request's arguments from the table and runs the body with its
// request-local index.  Results equal the per-request launches bit for bit: every workgroup
// runs the same code on the same data.
//
// Host side: vlaunch<F>(kernel, grid, block, shmem, stream, args...) launches the plain kernel,
// or, while a VRecorder is installed on the calling thread, appends the launch (group kernel,
// shape and packed arguments) to it instead.  Recording a request's step gives the launch
// sequence from which the work tables of a group are built.
#pragma once
#include "common.hpp"
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

struct VBlk {
    unsigned x, y, z;
};
// workgroup index (b) and grid size (g) as a kernel body sees them
struct VB {
    VBlk b, g;
};

__device__ __forceinline__ VB vb_hw() {
    return VB{{blockIdx.x, blockIdx.y, blockIdx.z}, {gridDim.x, gridDim.y, gridDim.z}};
}

// Argument pack with one layout on the host and the device (plain aggregate).
template <class... P> struct VPack;
template <> struct VPack<> {};
template <class H, class... T> struct VPack<H, T...> {
    H h;
    VPack<T...> t;
};

template <class... P> struct VMake;
template <> struct VMake<> {
    static VPack<> make() { return VPack<>{}; }
};
template <class H, class... T> struct VMake<H, T...> {
    static VPack<H, T...> make(H h, T... t) { return VPack<H, T...>{h, VMake<T...>::make(t...)}; }
};

template <size_t I, class H, class... T> __device__ __forceinline__ auto vget(const VPack<H, T...>& p) {
    if constexpr (I == 0) return p.h;
    else return vget<I - 1>(p.t);
}

// Work table of one grouped launch (device memory).  start[r] .. start[r+1] are request r's
// workgroups in the launch; grid[r] its own grid (x, y, z); its packed arguments sit at
// args + r * stride.  The argument rows are read through the constant address space, so the
// pointers they hold are known to be global memory (global, not flat, loads and stores).
struct VGroupTab {
    const u32* start;
    const VBlk* grid;
    u64 args;
    u32 n, stride;
};

#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) unsigned char* vcl_const_bytes;
#endif

// Request of this workgroup: the number of requests whose workgroups end at or before it
// (start is ascending), counted 64 entries per wave step.
__device__ __forceinline__ u32 vgroup_member(const VGroupTab& t) {
    const u32 bid = blockIdx.x;
    const u32 lane = __lane_id();
    u32 r = 0;
    for (u32 base = 0; base < t.n; base += 64) {
        const u32 i = base + lane;
        const bool done = i < t.n && t.start[i + 1] <= bid;
        const u32 c = (u32)__popcll(__ballot(done));
        r += c;
        if (c < 64) break;
    }
    return r;
}

template <class F, class... P, size_t... I>
__device__ __forceinline__ void vcall(const VB& vb, const VPack<P...>& p, std::index_sequence<I...>) {
    F{}(vb, vget<I>(p)...);
}

template <class F, class... P>
__global__ void __launch_bounds__(F::kBS, F::kMINW) vcl_group_kernel(VGroupTab t) {
    const u32 r = vgroup_member(t);
    const u32 local = blockIdx.x - t.start[r];
    const VBlk g = t.grid[r];
    const VB vb{{local % g.x, (local / g.x) % g.y, local / (g.x * g.y)}, g};
#if defined(__HIP_DEVICE_COMPILE__)
    const VPack<P...> p = *(const __attribute__((address_space(4))) VPack<P...>*)((vcl_const_bytes)t.args + (size_t)r * t.stride);
    vcall<F>(vb, p, std::index_sequence_for<P...>{});
#endif
}

// ----------------------------------------------------------------------------- host side
struct VLaunchRec {
    const void* fn;   // vcl_group_kernel<F, P...>
    dim3 grid, block;
    unsigned shmem;
    std::vector<unsigned char> args;   // VPack<P...>
};

struct VRecorder {
    std::vector<VLaunchRec> recs;
};

// Installed by the group builder while it records a request's step (one thread).
extern thread_local VRecorder* g_vcl_recorder;

template <class F, class... P, class... A>
inline void vlaunch(void (*kernel)(P...), dim3 grid, dim3 block, unsigned shmem, hipStream_t s, A&&... a) {
    if (g_vcl_recorder) {
        const VPack<P...> pk = VMake<P...>::make(static_cast<P>(a)...);
        VLaunchRec r;
        r.fn = (const void*)&vcl_group_kernel<F, P...>;
        r.grid = grid;
        r.block = block;
        r.shmem = shmem;
        r.args.resize(sizeof(pk));
        std::memcpy(r.args.data(), &pk, sizeof(pk));
        g_vcl_recorder->recs.push_back(std::move(r));
        return;
    }
    kernel<<<grid, block, shmem, s>>>(static_cast<P>(a)...);
}
