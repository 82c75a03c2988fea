// group.hip -- one step of several fit requests in shared launches.
//
// A group is built from the requests' recorded steps (vcl_fit_record_step): the requests must
// launch the same kernels in the same order with the same workgroup shape (same step
// signature).  For each launch position the group holds a device work table (group.hpp): the
// prefix of the requests' workgroup counts, their grids and their packed arguments.  A group
// step is then one launch per position, each covering every request's workgroups, so small
// requests fill the device together instead of leaving it idle between their own launches, and
// the launch count per step no longer grows with the number of requests.  The arguments are
// pointers into the requests' arenas and sizes, which do not change between steps (the
// iteration-dependent Adam state lives on the device), so the tables are built once per group.
#include "internal.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr size_t TAB_ALIGN = 256;
size_t al(size_t x) { return (x + TAB_ALIGN - 1) & ~(TAB_ALIGN - 1); }

struct Pos {
    const void* fn;
    dim3 block;
    unsigned shmem;
    u32 total;        // workgroups over all requests
    VGroupTab tab;    // device pointers into the table memory
};

// Sequences of every request, checked against the first; fills the per-position layout.
struct Built {
    std::vector<VRecorder> recs;
    std::vector<size_t> off_start, off_grid, off_args;
    std::vector<u32> stride;
    size_t bytes = 0;
};

int build(vcl_fit* const* fits, int n, Built& b) {
    VCL_REQUIRE(fits && n >= 1);
    b.recs.resize(n);
    for (int i = 0; i < n; i++) {
        VCL_REQUIRE(fits[i]);
        for (int j = 0; j < i; j++) VCL_REQUIRE(fits[j] != fits[i]);
        int st = vcl_fit_record_step(fits[i], &b.recs[i]);
        if (st) return st;
    }
    const auto& r0 = b.recs[0].recs;
    const size_t np = r0.size();
    for (int i = 1; i < n; i++) {
        const auto& ri = b.recs[i].recs;
        if (ri.size() != np) return VCL_ERR_INVALID;
        for (size_t k = 0; k < np; k++) {
            if (ri[k].fn != r0[k].fn || ri[k].shmem != r0[k].shmem || ri[k].args.size() != r0[k].args.size()) return VCL_ERR_INVALID;
            if (ri[k].block.x != r0[k].block.x || ri[k].block.y != r0[k].block.y || ri[k].block.z != r0[k].block.z) return VCL_ERR_INVALID;
        }
    }
    size_t o = 0;
    for (size_t k = 0; k < np; k++) {
        const u32 st = (u32)((r0[k].args.size() + 15) & ~(size_t)15);
        b.stride.push_back(st);
        b.off_start.push_back(o); o = al(o + sizeof(u32) * (n + 1));
        b.off_grid.push_back(o); o = al(o + sizeof(VBlk) * n);
        b.off_args.push_back(o); o = al(o + (size_t)st * n);
    }
    b.bytes = o ? o : TAB_ALIGN;
    return VCL_OK;
}

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}

}  // namespace

struct vcl_group {
    std::vector<vcl_fit*> fits;
    std::vector<Pos> pos;
    std::vector<unsigned char> host;   // the tables as uploaded (kept for the group's life)
    void* dev = nullptr;
    bool own = false;
    size_t bytes = 0;
    // HIP graph of one group step (vcl_group_step_graph)
    hipStream_t gstream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

extern "C" int vcl_fit_step_signature(vcl_fit* F, uint64_t* sig) {
    VCL_REQUIRE(F && sig);
    VRecorder rec;
    int st = vcl_fit_record_step(F, &rec);
    if (st) return st;
    uint64_t h = 1469598103934665603ull;
    for (const VLaunchRec& r : rec.recs) {
        h = fnv(h, &r.fn, sizeof(r.fn));
        const unsigned d[5] = {r.block.x, r.block.y, r.block.z, r.shmem, (unsigned)r.args.size()};
        h = fnv(h, d, sizeof(d));
    }
    *sig = h;
    return VCL_OK;
}

extern "C" size_t vcl_group_table_bytes(vcl_fit* const* fits, int n) {
    Built b;
    return build(fits, n, b) ? 0 : b.bytes;
}

extern "C" int vcl_group_create_in(vcl_fit* const* fits, int n, void* table, size_t table_bytes, vcl_stream_t stream, vcl_group** out) {
    VCL_REQUIRE(out && table);
    Built b;
    int st = build(fits, n, b);
    if (st) return st;
    if (table_bytes < b.bytes) return VCL_ERR_OOM;
    vcl_group* G = new vcl_group();
    G->fits.assign(fits, fits + n);
    G->host.assign(b.bytes, 0);
    G->dev = table;
    G->bytes = b.bytes;
    const size_t np = b.recs[0].recs.size();
    char* dev = (char*)table;
    for (size_t k = 0; k < np; k++) {
        u32* start = (u32*)(G->host.data() + b.off_start[k]);
        VBlk* grid = (VBlk*)(G->host.data() + b.off_grid[k]);
        unsigned char* args = G->host.data() + b.off_args[k];
        uint64_t total = 0;
        for (int i = 0; i < n; i++) {
            const VLaunchRec& r = b.recs[i].recs[k];
            start[i] = (u32)total;
            grid[i] = VBlk{r.grid.x, r.grid.y, r.grid.z};
            total += (uint64_t)r.grid.x * r.grid.y * r.grid.z;
            std::memcpy(args + (size_t)b.stride[k] * i, r.args.data(), r.args.size());
        }
        const VLaunchRec& r0 = b.recs[0].recs[k];
        const uint64_t threads = (uint64_t)r0.block.x * r0.block.y * r0.block.z;
        if (total == 0 || total > 0x7fffffffull || total * threads > 0xffffffffull) { delete G; return VCL_ERR_INVALID; }
        start[n] = (u32)total;
        Pos p;
        p.fn = r0.fn;
        p.block = r0.block;
        p.shmem = r0.shmem;
        p.total = (u32)total;
        p.tab.start = (const u32*)(dev + b.off_start[k]);
        p.tab.grid = (const VBlk*)(dev + b.off_grid[k]);
        p.tab.args = (u64)(uintptr_t)(dev + b.off_args[k]);
        p.tab.n = (u32)n;
        p.tab.stride = b.stride[k];
        G->pos.push_back(p);
    }
    if (hipMemcpyAsync(table, G->host.data(), b.bytes, hipMemcpyHostToDevice, as_hip(stream)) != hipSuccess) { delete G; return VCL_ERR_HIP; }
    *out = G;
    return VCL_OK;
}

extern "C" int vcl_group_create(vcl_fit* const* fits, int n, vcl_stream_t stream, vcl_group** out) {
    VCL_REQUIRE(out);
    size_t bytes = vcl_group_table_bytes(fits, n);
    if (!bytes) return VCL_ERR_INVALID;
    void* mem = nullptr;
    VCL_HIP_CHECK(hipMalloc(&mem, bytes));
    int st = vcl_group_create_in(fits, n, mem, bytes, stream, out);
    if (st) { (void)hipFree(mem); return st; }
    (*out)->own = true;
    return VCL_OK;
}

static int launch_all(vcl_group* G, hipStream_t s) {
    for (Pos& p : G->pos) {
        void* args[] = {&p.tab};
        VCL_HIP_CHECK(hipLaunchKernel(p.fn, dim3(p.total), p.block, args, p.shmem, s));
    }
    return VCL_OK;
}

extern "C" int vcl_group_step(vcl_group* G, vcl_stream_t stream) {
    VCL_REQUIRE(G);
    int st = launch_all(G, as_hip(stream));
    if (st) return st;
    for (vcl_fit* f : G->fits) vcl_fit_count_steps(f, 1);
    return VCL_OK;
}

extern "C" int vcl_group_step_graph(vcl_group* G, vcl_stream_t stream) {
    VCL_REQUIRE(G);
    if (!G->gexec) {
        VCL_HIP_CHECK(hipStreamCreateWithFlags(&G->gstream, hipStreamNonBlocking));
        VCL_HIP_CHECK(hipEventCreateWithFlags(&G->ev_in, hipEventDisableTiming));
        VCL_HIP_CHECK(hipEventCreateWithFlags(&G->ev_out, hipEventDisableTiming));
        VCL_HIP_CHECK(hipStreamBeginCapture(G->gstream, hipStreamCaptureModeThreadLocal));
        int st = launch_all(G, G->gstream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(G->gstream, &g);
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        VCL_HIP_CHECK(e);
        G->graph = g;
        VCL_HIP_CHECK(hipGraphInstantiate(&G->gexec, G->graph, nullptr, nullptr, 0));
    }
    hipStream_t hs = as_hip(stream);
    VCL_HIP_CHECK(hipEventRecord(G->ev_in, hs));
    VCL_HIP_CHECK(hipStreamWaitEvent(G->gstream, G->ev_in, 0));
    VCL_HIP_CHECK(hipGraphLaunch(G->gexec, G->gstream));
    VCL_HIP_CHECK(hipEventRecord(G->ev_out, G->gstream));
    VCL_HIP_CHECK(hipStreamWaitEvent(hs, G->ev_out, 0));
    for (vcl_fit* f : G->fits) vcl_fit_count_steps(f, 1);
    return VCL_OK;
}

extern "C" int vcl_group_size(const vcl_group* G) { return G ? (int)G->fits.size() : 0; }

extern "C" int vcl_group_launches(const vcl_group* G) { return G ? (int)G->pos.size() : 0; }

extern "C" int vcl_group_destroy(vcl_group* G) {
    if (!G) return VCL_OK;
    if (G->gstream) (void)hipStreamSynchronize(G->gstream);
    if (G->gexec) (void)hipGraphExecDestroy(G->gexec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    if (G->gstream) (void)hipStreamDestroy(G->gstream);
    if (G->ev_in) (void)hipEventDestroy(G->ev_in);
    if (G->ev_out) (void)hipEventDestroy(G->ev_out);
    if (G->own && G->dev) (void)hipFree(G->dev);
    delete G;
    return VCL_OK;
}
