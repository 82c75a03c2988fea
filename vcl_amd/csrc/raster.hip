// raster.hip -- tiled software rasterizer for gfx950 (replaces the OpenGL draw of
// Rasterize::forward, rasterize.cpp:257-282) and its backward (rasterize.cu:84-141).
//
// Pipeline per call (all on one stream, no host sync, no allocation):
//   1. clear      : zero tile counters, merge counters and the overflow counters
//   2. setup      : one thread per (layer, triangle).  Guard-band test, near/guard-band
//                   clipping when needed, 8-bit sub-pixel snapping, int64 edge equations
//                   with the top-left rule, depth planes of the original triangle; one
//                   112-byte record per piece.  A small piece (box of at most 4 tiles) counts
//                   itself into every tile whose pixel-centre lattice it can reach (exact
//                   edge test at the tile's extreme corner), aggregated per workgroup in an
//                   LDS hash table; larger pieces go to a list.
//   3. bin_large  : one wave per listed piece counts its tiles.
//   4. scan       : one workgroup: exclusive prefix of the tile counters (bin offsets) and the
//                   fine-raster work list (each bin cut into chunks of up to 256 pieces).
//   5. fill       : small pieces written into their tiles' bins (LDS-aggregated reservation)
//                   and, in the same launch, large pieces by one wave each.  A tile whose bin
//                   would pass the bin capacity scans every piece instead.
//   6. fine       : one 256-thread workgroup per chunk; each thread walks the exact covered
//                   row spans of one piece and takes the LDS minimum of the (ordered z_ndc,
//                   triangle id) key per pixel; a tile split over several chunks merges the
//                   partial keys in its last chunk.  The winner's (u, v, z/w, id+1) and DB are
//                   written with 16-byte stores.
// Coverage and depth use the exact float/int arithmetic of the CPU oracle's spec; this file
// is compiled without FP contraction so both produce identical triangle IDs.
#pragma clang fp contract(off)
#include "internal.hpp"
#include "raster_resolve.hpp"
#include "raster_span.hpp"
#include <algorithm>

namespace {

// Tunables (vcl_set_tuning): fine-raster tile side (8, 16 or 32 pixels) and pieces per
// fine-raster chunk (64, 128 or 256).  Defaults: 16 and 256.
constexpr int TILE_MAX = 32;
constexpr size_t BIN_ENTRIES_PER_PIECE = 32;   // bin capacity per piece (average over pieces)
constexpr size_t BIN_MIN_CAP = (size_t)1 << 22;
constexpr float RAST_GB = 16.0f;
constexpr float RAST_WEPS = 1e-6f;
constexpr u32 INVALID = 0xffffffffu;

struct __align__(16) RRec {
    int A[3];
    int B[3];
    int pxr;     // px0 | px1 << 16
    int pyr;     // py0 | py1 << 16
    long long C[3];
    u32 id;
    u32 view_tl; // view << 8 | top-left bits
    float zp[9]; // depth planes of the original triangle: Az Bz Cz Aw Bw Cw Ad Bd Cd
    float pad[3];
};
static_assert(sizeof(RRec) == 112, "record size");

constexpr int SMALL_TILES = 4;  // pieces whose box spans more tiles are binned by a whole wave
constexpr int LARGE_BLOCKS = 4096;   // waves for the large-piece list: about one piece each, as each costs two dependent loads

// Fine-raster work unit: up to `chunk` pieces of one tile's bin.  A tile whose bin holds more
// pieces is split over several chunks (and workgroups); the last of them to finish merges
// the partial depth keys.
constexpr u32 CHUNK_SCAN_ALL = 0x80000000u;   // count flag: scan every piece of every layer
constexpr int FINE_MAX_GRID = 65536;
// The large pieces of a tile are counted and placed through LSUB sub-counters (piece slot mod
// LSUB), laid out sub-counter-major so each sits in its own cache line: device atomics on one
// line serialise, and a deep tile takes ~1000 large pieces.
constexpr int LSUB = 16;
constexpr int CULL_MIN_BOX = 32;   // tile pixels in a piece's box for its depth range to be formed

struct Chunk {
    u32 tile, first, count, k;   // pieces [first, first + count) of the bin; k-th chunk of the tile
};

struct Layout {
    int tilesX, tilesY, ntiles, NT;
    size_t nrec_direct, ovf_cap, nrec_cap, bin_cap, chunk_cap;
    size_t off_counts, off_offsets, off_misc, off_rec, off_bins, off_large, off_chunks, off_keys, off_lcounts, off_lcursor, total;
};

size_t g_bin_cap_override = 0;   // vcl_rasterize_set_bin_capacity (0: sized from the piece count)

Layout make_layout(int W, int H, int depth, int T, int tshift) {
    Layout L;
    const int ts = 1 << tshift;
    L.tilesX = (W + ts - 1) / ts;
    L.tilesY = (H + ts - 1) / ts;
    L.ntiles = L.tilesX * L.tilesY;
    L.NT = L.ntiles * depth;
    L.nrec_direct = (size_t)depth * T;
    L.ovf_cap = L.nrec_direct / 2 + 1024;
    L.nrec_cap = L.nrec_direct + L.ovf_cap;
    L.bin_cap = L.nrec_cap * BIN_ENTRIES_PER_PIECE;
    if (L.bin_cap < BIN_MIN_CAP) L.bin_cap = BIN_MIN_CAP;
    if (g_bin_cap_override) L.bin_cap = g_bin_cap_override;
    L.chunk_cap = 2 * (size_t)L.NT + L.bin_cap / 64 + 1;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o = 0;
    L.off_counts = o; o = al(o + sizeof(u32) * (size_t)L.NT);
    L.off_offsets = o; o = al(o + sizeof(u32) * ((size_t)L.NT + 1));
    L.off_misc = o; o = al(o + sizeof(u32) * 64);   // [0]=overflow-piece count [2]=overflow flag [3]=large count [5]=chunks
    L.off_rec = o; o = al(o + sizeof(RRec) * L.nrec_cap);
    L.off_bins = o; o = al(o + sizeof(u32) * L.bin_cap);
    L.off_large = o; o = al(o + sizeof(u32) * L.nrec_cap);
    L.off_chunks = o; o = al(o + sizeof(Chunk) * L.chunk_cap);
    L.off_keys = o; o = al(o + sizeof(u64) * (size_t)W * H * depth);
    L.off_lcounts = o; o = al(o + sizeof(u32) * LSUB * (size_t)L.NT);
    L.off_lcursor = o; o = al(o + sizeof(u32) * LSUB * (size_t)L.NT);
    L.total = o;
    return L;
}

struct Ws {
    u32* counts; u32* offsets; u32* misc; RRec* rec; u32* bins; u32* large; Chunk* chunks; u64* keys;
    u32* lcounts;   // large pieces per (sub-counter, tile): [(slot % LSUB) * NT + tile]
    u32* lcursor;   // fill cursor of each (sub-counter, tile), starting at its range's first slot
};

Ws ws_ptrs(void* base, const Layout& L) {
    char* b = (char*)base;
    Ws w;
    w.counts = (u32*)(b + L.off_counts);
    w.offsets = (u32*)(b + L.off_offsets);
    w.misc = (u32*)(b + L.off_misc);
    w.rec = (RRec*)(b + L.off_rec);
    w.bins = (u32*)(b + L.off_bins);
    w.large = (u32*)(b + L.off_large);
    w.chunks = (Chunk*)(b + L.off_chunks);
    w.keys = (u64*)(b + L.off_keys);
    w.lcounts = (u32*)(b + L.off_lcounts);
    w.lcursor = (u32*)(b + L.off_lcursor);
    return w;
}

struct RastArgs {
    int W, H, depth, V, T, enableDB;
    float xs, ys;
    const float* proj;
    const u32* idx;
    float* out;
    float* outDB;
    int tilesX, ntiles, tshift;
    u32 chunk;
    size_t nrec_direct, ovf_cap, bin_cap, chunk_cap;
    Ws ws;
};

__device__ __forceinline__ u32 ord_f32(float z) {
    u32 b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ bool vtx_fast(float4 v) {
    return v.w > 0.0f && fabsf(v.x) <= RAST_GB * v.w && fabsf(v.y) <= RAST_GB * v.w;
}

// Snap + edge setup of one piece; false if it has zero area or no pixel centre in its box.
__device__ bool setup_piece(float4 a, float4 b, float4 c, u32 id, u32 view, int W, int H, RRec& r) {
    float4 v[3] = {a, b, c};
    int X[3], Y[3];
    float halfW = 0.5f * (float)W, halfH = 0.5f * (float)H;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float nx = v[i].x / v[i].w, ny = v[i].y / v[i].w;
        float fx = (nx + 1.0f) * halfW, fy = (ny + 1.0f) * halfH;
        X[i] = (int)rintf(fx * 256.0f);
        Y[i] = (int)rintf(fy * 256.0f);
    }
    i64 area = (i64)(X[1] - X[0]) * (Y[2] - Y[0]) - (i64)(Y[1] - Y[0]) * (X[2] - X[0]);
    if (area == 0) return false;
    i64 s = area > 0 ? 1 : -1;
    u32 tl = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int j = (i + 1) % 3;
        i64 dx = X[j] - X[i], dy = Y[j] - Y[i];
        i64 A = -s * dy, B = s * dx;
        r.A[i] = (int)A;
        r.B[i] = (int)B;
        r.C[i] = s * (dy * X[i] - dx * Y[i]);
        if (A > 0 || (A == 0 && B < 0)) tl |= 1u << i;
    }
    int xmin = min(X[0], min(X[1], X[2])), xmax = max(X[0], max(X[1], X[2]));
    int ymin = min(Y[0], min(Y[1], Y[2])), ymax = max(Y[0], max(Y[1], Y[2]));
    int px0 = max(0, (xmin - 128 + 255) >> 8), px1 = min(W - 1, (xmax - 128) >> 8);
    int py0 = max(0, (ymin - 128 + 255) >> 8), py1 = min(H - 1, (ymax - 128) >> 8);
    if (px0 > px1 || py0 > py1) return false;
    r.pxr = px0 | (px1 << 16);
    r.pyr = py0 | (py1 << 16);
    r.id = id;
    r.view_tl = (view << 8) | tl;
    return true;
}

__device__ int clip_poly(const float4* in, int n, float4* out, float p0, float p1, float p2, float p3, float p4) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        float4 a = in[i], b = in[(i + 1) % n];
        float da = p0 * a.x + p1 * a.y + p2 * a.z + p3 * a.w + p4;
        float db = p0 * b.x + p1 * b.y + p2 * b.z + p3 * b.w + p4;
        if (da >= 0.0f) out[m++] = a;
        if ((da >= 0.0f) != (db >= 0.0f)) {
            float t = da / (da - db);
            out[m++] = make_float4(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w));
        }
    }
    return m;
}

// Can any pixel centre of tile (tx, ty) be inside the piece?  Each edge function is maximised
// over the tile's centre lattice at one corner; the piece misses the tile when that maximum
// fails the inside test (E > 0, or E == 0 on a top-left edge).
__device__ __forceinline__ bool tile_overlaps(const RRec& r, int tx, int ty, int tshift) {
    long long x0 = (long long)(tx << tshift) * 256 + 128, x1 = x0 + (long long)((1 << tshift) - 1) * 256;
    long long y0 = (long long)(ty << tshift) * 256 + 128, y1 = y0 + (long long)((1 << tshift) - 1) * 256;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        long long X = r.A[i] > 0 ? x1 : x0, Y = r.B[i] > 0 ? y1 : y0;
        long long e = (long long)r.A[i] * X + (long long)r.B[i] * Y + r.C[i];
        if (e < 0 || (e == 0 && !((r.view_tl >> i) & 1u))) return false;
    }
    return true;
}

__device__ __forceinline__ int box_tiles(const RRec& r, int tshift) {
    int tx0 = (r.pxr & 0xffff) >> tshift, tx1 = (r.pxr >> 16) >> tshift;
    int ty0 = (r.pyr & 0xffff) >> tshift, ty1 = (r.pyr >> 16) >> tshift;
    return (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
}

// ---------------------------------------------------------------- binning
// The per-tile counters are hot (hundreds of pieces land on one tile), so every binning pass
// first aggregates in an LDS hash table keyed by the global tile index (view * ntiles + tile):
// one global atomic per distinct tile per workgroup.  A key that finds no room in the table
// falls back to a direct global atomic.
constexpr int BIN_SLOTS = 2048;

// key -> val.  Count pass: val = entries.  Fill pass: val = entries, then (bin_fill_reserve)
// the first position of the workgroup's run in the tile's bin, advanced by each write.
struct BinTable {
    u32 key[BIN_SLOTS];
    u32 val[BIN_SLOTS];
};

__device__ __forceinline__ void bin_table_clear(BinTable& t) {
    for (int i = threadIdx.x; i < BIN_SLOTS; i += blockDim.x) { t.key[i] = LDS_EMPTY; t.val[i] = 0; }
}

__device__ __forceinline__ int bin_slot(BinTable& t, u32 k) {
    u32 h = (k * 2654435761u) & (BIN_SLOTS - 1);
    for (int probe = 0; probe < 32; probe++) {
        u32 old = atomicCAS(&t.key[h], LDS_EMPTY, k);
        if (old == LDS_EMPTY || old == k) return (int)h;
        h = (h + 1) & (BIN_SLOTS - 1);
    }
    return -1;
}

// count pass: one piece-tile entry
__device__ __forceinline__ void bin_count(const RastArgs& a, BinTable& t, u32 g) {
    int sl = bin_slot(t, g);
    if (sl >= 0) atomicAdd(&t.val[sl], 1u);
    else atomicAdd(a.ws.counts + g, 1u);
}

// count pass: flush the workgroup's per-tile totals (n entries of (key, val))
__device__ __forceinline__ void bin_count_flush(const RastArgs& a, const u32* key, const u32* val, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (val[i] && (!key || key[i] != LDS_EMPTY)) atomicAdd(a.ws.counts + (key ? key[i] : (u32)i), val[i]);
}

// fill pass: reserve the workgroup's run of each tile's bin (the counters count down from the
// totals of the count pass); val becomes the first position of the run within the tile's bin
__device__ __forceinline__ void bin_fill_reserve(const RastArgs& a, const u32* key, u32* val, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (val[i] && (!key || key[i] != LDS_EMPTY)) val[i] = atomicSub(a.ws.counts + (key ? key[i] : (u32)i), val[i]) - val[i];
}

__device__ __forceinline__ void bin_write(const RastArgs& a, u32 g, u32 pos, u32 slot) {
    size_t at = (size_t)a.ws.offsets[g] + pos;
    if (at < a.bin_cap) a.ws.bins[at] = slot;
}

// fill pass, direct path for an entry whose key found no table slot
__device__ __forceinline__ void bin_fill_direct(const RastArgs& a, u32 g, u32 slot) {
    bin_write(a, g, atomicSub(a.ws.counts + g, 1u) - 1u, slot);
}

__device__ __forceinline__ bool piece_small(const RRec& r, int tshift) { return box_tiles(r, tshift) <= SMALL_TILES; }

// Tiles (global index) of a small piece that its pixel-centre lattice can reach; returns the count.
__device__ __forceinline__ int small_tiles(const RastArgs& a, const RRec& r, u32* g) {
    int view = (int)(r.view_tl >> 8);
    int tx0 = (r.pxr & 0xffff) >> a.tshift, tx1 = (r.pxr >> 16) >> a.tshift;
    int ty0 = (r.pyr & 0xffff) >> a.tshift, ty1 = (r.pyr >> 16) >> a.tshift;
    bool single = tx0 == tx1 && ty0 == ty1;
    int n = 0;
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
            if (!single && !tile_overlaps(r, tx, ty, a.tshift)) continue;
            if (n < SMALL_TILES) g[n++] = (u32)(view * a.ntiles + ty * a.tilesX + tx);
        }
    return n;
}

// Large pieces (box over SMALL_TILES tiles): one wave per piece, lanes walk the piece's tiles,
// so the counter atomics of a long sliver overlap instead of running one after another in a
// single lane.  These entries spread over many tiles per workgroup, so they go straight to
// the global counters (LDS aggregation measured slower here).
template <bool FILL>
struct KBinLargeF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a) const {
        const int lane = threadIdx.x & 63;
        const int wave = (vb.b.x * blockDim.x + threadIdx.x) >> 6;
        const int nwaves = (vb.g.x * blockDim.x) >> 6;
        const u32 nlarge = a.ws.misc[3];
        for (u32 i = wave; i < nlarge; i += nwaves) {
            u32 slot = a.ws.large[i];
            const RRec& r = a.ws.rec[slot];
            int view = (int)(r.view_tl >> 8);
            int tx0 = (r.pxr & 0xffff) >> a.tshift, tx1 = (r.pxr >> 16) >> a.tshift;
            int ty0 = (r.pyr & 0xffff) >> a.tshift, ty1 = (r.pyr >> 16) >> a.tshift;
            int ntx = tx1 - tx0 + 1, n = ntx * (ty1 - ty0 + 1);
            for (int k = lane; k < n; k += 64) {
                int tx = tx0 + k % ntx, ty = ty0 + k / ntx;
                if (!tile_overlaps(r, tx, ty, a.tshift)) continue;
                u32 g = (u32)(view * a.ntiles + ty * a.tilesX + tx);
                const u32 sub = slot & (LSUB - 1);
                const size_t NT = (size_t)a.ntiles * a.depth;
                if (!FILL) {
                    atomicAdd(a.ws.lcounts + sub * NT + g, 1u);
                } else {
                    // the tile's large pieces take the bottom of its bin (the small ones count down
                    // from the top): sub-counter ranges in order, each filled by its cursor
                    bin_write(a, g, atomicAdd(a.ws.lcursor + sub * NT + g, 1u), slot);
                }
            }
        }
    }
};
template <bool FILL>
__global__ void __launch_bounds__(256) k_bin_large(RastArgs a) { KBinLargeF<FILL>{}(vb_hw(), a); }

// Workgroup-local list of the large pieces found by setup, appended to the global list with
// one atomic per workgroup.
constexpr int SETUP_LQ = 512;

struct SetupShared {
    BinTable tab;
    u32 lq[SETUP_LQ];
    u32 nlq, lbase;
};

__device__ __forceinline__ void bin_or_defer(const RastArgs& a, SetupShared& sh, const RRec& r, u32 slot) {
    if (piece_small(r, a.tshift)) {
        u32 g[SMALL_TILES];
        int n = small_tiles(a, r, g);
        for (int i = 0; i < n; i++) bin_count(a, sh.tab, g[i]);
    } else {
        u32 i = atomicAdd(&sh.nlq, 1u);
        if (i < SETUP_LQ) sh.lq[i] = slot;
        else a.ws.large[atomicAdd(a.ws.misc + 3, 1u)] = slot;
    }
}

// Per-launch workspace reset: tile counters, large-piece sub-counters and misc[] to zero, the
// per-pixel depth keys to all-ones (no winner); grid-stride over the largest.
struct KClearF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, WsClear c) const {
        const size_t n = c.n();
        for (size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x; i < n; i += (size_t)vb.g.x * blockDim.x) ws_clear_elem(c, i);
    }
};
__global__ void __launch_bounds__(256) k_clear(WsClear c) { KClearF{}(vb_hw(), c); }

WsClear ws_clear_of(const Ws& w, const Layout& L, size_t P) {
    WsClear c;
    c.counts = w.counts; c.nt = L.NT;
    c.lcounts = w.lcounts; c.nl = (size_t)LSUB * L.NT;
    c.misc = w.misc;
    c.keys = (unsigned long long*)w.keys; c.np = P;
    return c;
}

// Adds each tile's large-piece sub-counts to its count, before the scan, and sets each
// sub-counter's fill cursor to the start of its range (the exclusive prefix of the sub-counts).
struct KFoldLargeF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, Ws w, int NT) const {
        const int g = vb.b.x * blockDim.x + threadIdx.x;
        if (g >= NT) return;
        u32 t = 0;
#pragma unroll
        for (int q = 0; q < LSUB; q++) {
            const size_t i = (size_t)q * NT + g;
            const u32 c = w.lcounts[i];
            w.lcursor[i] = t;
            t += c;
        }
        if (t) w.counts[g] += t;
    }
};
__global__ void __launch_bounds__(256) k_fold_large(Ws w, int NT) { KFoldLargeF{}(vb_hw(), w, NT); }

__device__ void setup_triangle(const RastArgs& a, SetupShared& sh, size_t gid) {
    if (gid >= a.nrec_direct) return;
    u32 view = (u32)(gid / a.T);
    u32 t = (u32)(gid % a.T);
    const float4* clip = (const float4*)a.proj + (size_t)view * a.V;
    u32 i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
    float4 v0 = clip[i0], v1 = clip[i1], v2 = clip[i2];
    RRec r;
    if (vtx_fast(v0) && vtx_fast(v1) && vtx_fast(v2)) {
        if (setup_piece(v0, v1, v2, t, view, a.W, a.H, r)) {
            depth_planes(v0, v1, v2, r.zp);
            a.ws.rec[gid] = r;
            bin_or_defer(a, sh, r, (u32)gid);
        } else {
            a.ws.rec[gid].id = INVALID;
        }
        return;
    }
    a.ws.rec[gid].id = INVALID;
    // clip path: near (z + w >= 0), w >= eps, guard band |x|,|y| <= GB*w
    float4 A[12], B[12];
    A[0] = v0; A[1] = v1; A[2] = v2;
    int n = 3;
    const float planes[6][5] = {{0, 0, 1, 1, 0}, {0, 0, 0, 1, -RAST_WEPS}, {-1, 0, 0, RAST_GB, 0},
                                {1, 0, 0, RAST_GB, 0}, {0, -1, 0, RAST_GB, 0}, {0, 1, 0, RAST_GB, 0}};
    for (int p = 0; p < 6 && n > 0; p++) {
        n = clip_poly(A, n, B, planes[p][0], planes[p][1], planes[p][2], planes[p][3], planes[p][4]);
        for (int i = 0; i < n; i++) A[i] = B[i];
    }
    float zp[9];
    depth_planes(v0, v1, v2, zp);
    for (int i = 1; i + 1 < n; i++) {
        if (!setup_piece(A[0], A[i], A[i + 1], t, view, a.W, a.H, r)) continue;
        for (int k = 0; k < 9; k++) r.zp[k] = zp[k];
        u32 s = atomicAdd(a.ws.misc + 0, 1u);
        if (s >= a.ovf_cap) { atomicOr(a.ws.misc + 2, 1u); continue; }
        u32 slot = (u32)(a.nrec_direct + s);
        a.ws.rec[slot] = r;
        bin_or_defer(a, sh, r, slot);
    }
}

// One thread per (layer, triangle): guard-band test, clipping when needed, snapping and edge
// setup, one record per piece; small pieces are counted into their tiles' bins, large pieces
// are listed for the wave-per-piece pass.
struct KSetupF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a) const {
        __shared__ SetupShared sh;
        bin_table_clear(sh.tab);
        if (threadIdx.x == 0) sh.nlq = 0;
        __syncthreads();
        setup_triangle(a, sh, (size_t)vb.b.x * blockDim.x + threadIdx.x);
        __syncthreads();
        bin_count_flush(a, sh.tab.key, sh.tab.val, BIN_SLOTS);
        u32 nl = min(sh.nlq, (u32)SETUP_LQ);
        if (threadIdx.x == 0 && nl) sh.lbase = atomicAdd(a.ws.misc + 3, nl);
        __syncthreads();
        for (u32 i = threadIdx.x; i < nl; i += blockDim.x) a.ws.large[sh.lbase + i] = sh.lq[i];
    }
};
__global__ void __launch_bounds__(256) k_setup(RastArgs a) { KSetupF{}(vb_hw(), a); }

// Exclusive prefix of one value per thread over a 1024-lane workgroup, continuing from *carry
// (which is advanced by the workgroup's total).
__device__ u32 scan1024(u32 v, u32* wsum, u32* carry) {
    int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    u32 x = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        u32 s = wave_incl_scan(lane < 16 ? wsum[lane] : 0u);
        if (lane < 16) wsum[lane] = s;
    }
    __syncthreads();
    u32 pre = *carry + (wid > 0 ? wsum[wid - 1] : 0u) + x - v;
    __syncthreads();
    if (tid == 1023) *carry = pre + v;
    __syncthreads();
    return pre;
}

// Single workgroup: exclusive scan of the NT tile counters into bin offsets, then the fine-raster
// work list: each tile's bin cut into chunks of `chunk` pieces (one chunk for an empty tile,
// and one scan-everything chunk for a tile whose bin would pass the bin capacity).
struct KScanF {
    static constexpr int kBS = 1024, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a, int NT) const {
        __shared__ u32 wsum[16];
        __shared__ u32 carry[2];
        int tid = threadIdx.x;
        if (tid == 0) { carry[0] = 0; carry[1] = 0; }
        __syncthreads();
        const u32 nrec = (u32)(a.nrec_direct + min(a.ws.misc[0], (u32)a.ovf_cap));
        // each thread takes TPT consecutive tiles: their local prefix, then one block scan of the
        // threads' totals (TPT times fewer block scans and barriers than one tile per thread)
        constexpr int TPT = 4;
        for (int base = 0; base < NT; base += 1024 * TPT) {
            const int i0 = base + TPT * tid;
            u32 v[TPT], loc[TPT], t = 0;
#pragma unroll
            for (int j = 0; j < TPT; j++) {
                v[j] = i0 + j < NT ? a.ws.counts[i0 + j] : 0u;
                loc[j] = t;
                t += v[j];
            }
            const u32 pre0 = scan1024(t, wsum, &carry[0]);
            u32 nch[TPT], cloc[TPT], tc = 0;
            bool all[TPT];
#pragma unroll
            for (int j = 0; j < TPT; j++) {
                const int i = i0 + j;
                const u32 pre = pre0 + loc[j];
                if (i < NT) a.ws.offsets[i] = pre;
                all[j] = (size_t)pre + v[j] > a.bin_cap;
                nch[j] = i >= NT ? 0u : all[j] ? 1u : max(1u, (v[j] + a.chunk - 1) / a.chunk);
                cloc[j] = tc;
                tc += nch[j];
            }
            const u32 cpre0 = scan1024(tc, wsum, &carry[1]);
#pragma unroll
            for (int j = 0; j < TPT; j++) {
                const u32 pre = pre0 + loc[j];
                for (u32 k = 0; k < nch[j]; k++) {
                    size_t c = (size_t)cpre0 + cloc[j] + k;
                    if (c >= a.chunk_cap) break;
                    Chunk ch;
                    ch.tile = (u32)(i0 + j);
                    ch.first = all[j] ? 0u : pre + k * a.chunk;
                    ch.count = all[j] ? (nrec | CHUNK_SCAN_ALL) : min(a.chunk, v[j] - k * a.chunk);
                    ch.k = k;
                    a.ws.chunks[c] = ch;
                }
            }
        }
        if (tid == 0) {
            a.ws.offsets[NT] = carry[0];
            a.ws.misc[5] = (u32)min((size_t)carry[1], a.chunk_cap);
        }
    }
};
__global__ void __launch_bounds__(1024) k_scan(RastArgs a, int NT) { KScanF{}(vb_hw(), a, NT); }

// Small pieces into their tiles' bins: entries counted into LDS (local position from the LDS
// atomic), each tile's run reserved with one global atomic, then written.
struct KFillF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a, size_t nslots) const {
        __shared__ BinTable tab;
        bin_table_clear(tab);
        __syncthreads();
        size_t slot = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        u32 g[SMALL_TILES], loc[SMALL_TILES];
        int sl[SMALL_TILES];
        int n = 0;
        bool ok = slot < nslots;
        if (ok && slot >= a.nrec_direct) {
            u32 novf = min(a.ws.misc[0], (u32)a.ovf_cap);
            ok = slot - a.nrec_direct < novf;
        }
        if (ok) {
            const RRec& r = a.ws.rec[slot];
            if (r.id != INVALID) {
                RRec rr = r;
                if (piece_small(rr, a.tshift)) n = small_tiles(a, rr, g);
            }
        }
        for (int i = 0; i < n; i++) {
            sl[i] = bin_slot(tab, g[i]);
            if (sl[i] >= 0) loc[i] = atomicAdd(&tab.val[sl[i]], 1u);
            else bin_fill_direct(a, g[i], (u32)slot);
        }
        __syncthreads();
        bin_fill_reserve(a, tab.key, tab.val, BIN_SLOTS);
        __syncthreads();
        for (int i = 0; i < n; i++)
            if (sl[i] >= 0) bin_write(a, g[i], tab.val[sl[i]] + loc[i], (u32)slot);
    }
};

// The small- and large-piece bin fills in one launch: workgroups [0, nfill) run the small-piece
// fill, the others the large-piece fill.  They write disjoint parts of each bin (the large pieces
// its bottom through the sub-counter cursors, the small ones its top, counting down), so the two
// overlap on the device instead of one waiting for the other's last workgroups.
struct KFillAllF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a, size_t nslots, u32 nfill) const {
        if (vb.b.x < nfill) KFillF{}(VB{vb.b, {nfill, vb.g.y, vb.g.z}}, a, nslots);
        else KBinLargeF<true>{}(VB{{vb.b.x - nfill, vb.b.y, vb.b.z}, {vb.g.x - nfill, vb.g.y, vb.g.z}}, a);
    }
};
__global__ void __launch_bounds__(256) k_fill_all(RastArgs a, size_t nslots, u32 nfill) { KFillAllF{}(vb_hw(), a, nslots, nfill); }

constexpr u32 SPAN_EMPTY = 0x00ffu;   // first column 255 > last column 0

// Winner of one pixel -> (u, v, z/w, id+1) and DB, written with 16-byte stores.
__device__ void fine_resolve(const RastArgs& a, u32 view, int px, int py, float pixx, float pixy, u64 best) {
    if (px >= a.W || py >= a.H) return;
    const float4* clip = (const float4*)a.proj + (size_t)view * a.V;
    size_t p = ((size_t)view * a.H + py) * a.W + px;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
    if (best != RAST_NO_KEY) {
        u32 t = (u32)(best & 0xffffffffu);
        resolve_pixel(clip[a.idx[3 * t]], clip[a.idx[3 * t + 1]], clip[a.idx[3 * t + 2]], t, a.xs, a.ys, pixx, pixy, a.outDB != nullptr, o, d);
    }
    ((float4*)a.out)[p] = o;
    if (a.outDB) ((float4*)a.outDB)[p] = d;
}

// ---- tile-level occlusion culling in the fine raster
// The per-pixel depth is z = Zp / Wp with Zp = (z0 x + z1 y) + z2 (float, uncontracted) and the
// same for Wp and Dp.  Over a rectangle of pixel centres on which the exact Wp keeps one sign,
// the exact z (linear-fractional) takes its extremes at the corners.  Any float evaluation of
// the same expressions (per pixel, or here at the corners) differs from the exact z by at most
// d = (eZ + |z| eW) / (|Wp| - 2 eW) + ulp(z), with |Wp| the smallest float |Wp| at the corners
// and eZ = 2 * 2^-23 * (|z0 x| + |z1 y| + |z2|) bounding the four roundings of Zp (likewise eW,
// eD); the corner values therefore bound every pixel's float z within 2 d, and eZ, eW are taken
// 4x larger to cover the float evaluation of the bound itself.  A
// piece that covers every pixel centre of the tile and passes every pixel test there writes a
// key of depth <= its upper bound at every pixel, so a piece whose lower bound is above that
// can never win a pixel and is skipped.

// Conservative [lo, hi] of the float per-pixel z of a piece over the rectangle with pixel-centre
// corners (xa, ya), (xb, yb); false when Wp may vanish in it or a value is not finite.  valid:
// every pixel of the rectangle passes the Wp / Dp sign test and -1 <= z <= 1.
__device__ __forceinline__ bool piece_depth_range(const float* zp, float xa, float xb, float ya, float yb, float& lo, float& hi, bool& valid) {
    float zmin = INFINITY, zmax = -INFINITY, wmin = INFINITY, dmin = INFINITY, SZ = 0.f, SW = 0.f, SD = 0.f;
    int wpos = 0, wneg = 0, dpos = 0, dneg = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float x = (c & 1) ? xb : xa, y = (c & 2) ? yb : ya;
        const float z0 = zp[0] * x, z1 = zp[1] * y, w0 = zp[3] * x, w1 = zp[4] * y, d0 = zp[6] * x, d1 = zp[7] * y;
        const float Z = z0 + z1 + zp[2], Wv = w0 + w1 + zp[5], Dv = d0 + d1 + zp[8];
        SZ = fmaxf(SZ, fabsf(z0) + fabsf(z1) + fabsf(zp[2]));
        SW = fmaxf(SW, fabsf(w0) + fabsf(w1) + fabsf(zp[5]));
        SD = fmaxf(SD, fabsf(d0) + fabsf(d1) + fabsf(zp[8]));
        wpos += Wv > 0.f; wneg += Wv < 0.f; dpos += Dv > 0.f; dneg += Dv < 0.f;
        wmin = fminf(wmin, fabsf(Wv)); dmin = fminf(dmin, fabsf(Dv));
        const float z = Z / Wv;
        zmin = fminf(zmin, z); zmax = fmaxf(zmax, z);
    }
    if (!(isfinite(SZ) && isfinite(SW) && isfinite(SD))) return false;
    const float U = 8.0f * 0x1p-23f;
    const float eZ = U * SZ + 1e-30f, eW = U * SW + 1e-30f, eD = U * SD + 1e-30f;
    if (!(wpos == 4 || wneg == 4) || !(wmin > 4.f * eW) || !isfinite(zmin) || !isfinite(zmax)) return false;
    const float za = fmaxf(fabsf(zmin), fabsf(zmax));
    const float d = 2.f * ((eZ + za * eW) / (wmin - 2.f * eW) + 0x1p-21f * (za + 1.f));
    lo = zmin - 1.01f * d;
    hi = zmax + 1.01f * d;
    valid = ((wpos == 4 && dpos == 4) || (wneg == 4 && dneg == 4)) && dmin > eD && lo >= -1.0f && hi <= 1.0f;
    return true;
}

// The piece covers every pixel centre of [px0, px1] x [py0, py1]: its integer edge functions
// (with the top-left offset, inside <=> > 0) are positive at the four corners, so everywhere.
__device__ __forceinline__ bool piece_covers_rect(const RRec& r, int px0, int px1, int py0, int py1) {
    const long long X0 = (long long)px0 * 256 + 128, X1 = (long long)px1 * 256 + 128;
    const long long Y0 = (long long)py0 * 256 + 128, Y1 = (long long)py1 * 256 + 128;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const long long c = (long long)r.C[i] + ((r.view_tl >> i) & 1u);
        const long long ax0 = (long long)r.A[i] * X0, ax1 = (long long)r.A[i] * X1;
        const long long by0 = (long long)r.B[i] * Y0, by1 = (long long)r.B[i] * Y1;
        if (ax0 + by0 + c <= 0 || ax1 + by0 + c <= 0 || ax0 + by1 + c <= 0 || ax1 + by1 + c <= 0) return false;
    }
    return true;
}

// One 256-thread workgroup per work-list chunk (up to `chunk` pieces of one TS x TS tile's bin,
// so a deep bin is spread over several workgroups).  Per batch of pieces: (1a) thread = piece
// forms the piece's conservative depth range over the tile, and pieces covering the whole tile
// lower the tile's depth bound; (1b) the pieces not behind that bound, compacted onto consecutive
// threads, find their exact covered column span in each row (a double-precision crossing
// estimate settled by exact edge tests, so no per-pixel edge test);
// (2) a block prefix turns the covered rows into items; (3) thread = item evaluates the depth
// planes at each covered pixel centre and takes the LDS 64-bit minimum of the (ordered z_ndc,
// triangle id) key.  The result does not depend on bin order, tile size, culling or how the
// bin is split: a tile's only chunk stores its keys, the chunks of a split tile merge theirs
// with a device-scope 64-bit minimum.
template <int TS>
struct KFineF {
    static constexpr int kBS = 256, kMINW = TS == 32 ? 4 : 6;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a) const {
        constexpr int NP = TS * TS;
        __shared__ u64 key[NP];
        __shared__ float spix[2][TS];
        constexpr int BATCH = TS == 32 ? 128 : 256;    // pieces per LDS batch
        __shared__ unsigned short sspan[BATCH][TS];     // per piece: first | last << 8 column per row
        __shared__ float szp[9][BATCH];
        __shared__ u32 sid[BATCH];
        constexpr int ITEM_CAP = BATCH * TS < 1024 ? BATCH * TS : 1024;
        __shared__ u32 sitem[ITEM_CAP];                  // covered rows: piece | row | first column | length
        __shared__ u32 sbucket[4];                       // per-wave covered-row totals
        __shared__ u32 swcnt[4];                         // per-wave surviving-piece counts
        __shared__ u32 sslot[BATCH];                     // surviving pieces' record slots
        __shared__ u32 sbound;                           // ordered depth bound of the tile (culling)
        const int tid = threadIdx.x;
        const u32 nchunks = a.ws.misc[5];
        for (u32 c = vb.b.x; c < nchunks; c += vb.g.x) {
            const Chunk ch = a.ws.chunks[c];
            const int tile = (int)ch.tile;
            const u32 view = (u32)(tile / a.ntiles);
            const int tl = tile % a.ntiles;
            const int tx = tl % a.tilesX, ty = tl / a.tilesX;
            for (int i = tid; i < NP; i += 256) key[i] = ~0ull;
            if (tid == 0) sbound = 0xffffffffu;
            if (tid < 2 * TS) {
                int cc = tid % TS;
                spix[tid / TS][cc] = tid < TS ? (float)((tx * TS + cc + 0.5) * (double)a.xs - 1.0)
                                              : (float)((ty * TS + cc + 0.5) * (double)a.ys - 1.0);
            }
            const int x0 = tx * TS, y0 = ty * TS;
            const long long X0 = (long long)x0 * 256 + 128;
            const bool scan_all = (ch.count & CHUNK_SCAN_ALL) != 0u;
            const u32 n = ch.count & ~CHUNK_SCAN_ALL;
            __syncthreads();
            // the tile's pixel rectangle inside the image, for the culling bounds
            const int rx1 = min(x0 + TS, a.W) - 1 - x0, ry1 = min(y0 + TS, a.H) - 1 - y0;
            for (u32 base = 0; base < n; base += BATCH) {
                // phase 1a: thread = piece; conservative depth range over the tile; pieces covering
                // the whole tile with every pixel test passing lower the tile's depth bound
                const u32 q = base + tid;
                int yl = 1, yh = 0;
                u32 nrow = 0;
                bool live = false;
                float zlo = 0.f;
                bool has_lo = false;
                u32 slot = 0;
                if (tid < BATCH && q < n) {
                    slot = scan_all ? q : a.ws.bins[ch.first + q];
                    const RRec r = a.ws.rec[slot];
                    int xl = max((r.pxr & 0xffff) - x0, 0), xh = min((r.pxr >> 16) - x0, TS - 1);
                    int yl0 = max((r.pyr & 0xffff) - y0, 0), yh0 = min((r.pyr >> 16) - y0, TS - 1);
                    live = r.id != INVALID && (r.view_tl >> 8) == view && xl <= xh && yl0 <= yh0;
                    // a piece whose box holds few of the tile's pixels saves little when culled and
                    // cannot cover the tile, so its range is not formed
                    if (live && (xh - xl + 1) * (yh0 - yl0 + 1) >= CULL_MIN_BOX) {
                        float hi;
                        bool valid;
                        has_lo = piece_depth_range(r.zp, spix[0][0], spix[0][rx1], spix[1][0], spix[1][ry1], zlo, hi, valid);
                        if (has_lo && valid && piece_covers_rect(r, x0, x0 + rx1, y0, y0 + ry1)) atomicMin(&sbound, ord_f32(hi));
                    }
                }
                __syncthreads();
                // the pieces not behind the bound, compacted in piece order, so that the span pass
                // below runs on consecutive threads (most pieces of a deep tile are culled)
                const bool surv = live && !(has_lo && ord_f32(zlo) > sbound);
                const u64 sm = __ballot(surv);
                if ((tid & 63) == 0) swcnt[tid >> 6] = (u32)__popcll(sm);
                __syncthreads();
                u32 spos = (u32)__popcll(sm & ((1ull << (tid & 63)) - 1ull));
                for (int w = 0; w < (tid >> 6); w++) spos += swcnt[w];
                const u32 nsurv = swcnt[0] + swcnt[1] + swcnt[2] + swcnt[3];
                if (surv) sslot[spos] = slot;
                __syncthreads();
                // phase 1b: thread j < nsurv takes the j-th surviving piece: exact row spans and depth
                // planes into LDS; covered rows counted
                if ((u32)tid < nsurv) {
                    const RRec r = a.ws.rec[sslot[tid]];
                    u32 id = r.id, vtl = r.view_tl;
                    int xl = max((r.pxr & 0xffff) - x0, 0), xh = min((r.pxr >> 16) - x0, TS - 1);
                    yl = max((r.pyr & 0xffff) - y0, 0);
                    yh = min((r.pyr >> 16) - y0, TS - 1);
                    {
                        // exact spans (row_span_exact): the edge functions at the first row, their
                        // steps per column and row, and the crossings' estimates -K / dA (first row)
                        // and their change per row
                        SpanEdges e;
                        double dB[3], t0[3], ts[3];
                        const long long Y = (long long)(y0 + yl) * 256 + 128;
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            e.K[i] = (double)((long long)r.A[i] * X0 + (long long)r.B[i] * Y + r.C[i] + ((vtl >> i) & 1u));
                            e.dA[i] = (double)r.A[i] * 256.0;
                            dB[i] = (double)r.B[i] * 256.0;
                            const double rA = e.dA[i] != 0.0 ? 1.0 / e.dA[i] : 0.0;
                            t0[i] = -e.K[i] * rA;
                            ts[i] = -dB[i] * rA;
                        }
                        for (int y = yl; y <= yh; y++) {
                            int lo = xl, hi = xh;
#pragma unroll
                            for (int i = 0; i < 3; i++) e.t[i] = fma(ts[i], (double)(y - yl), t0[i]);
                            row_span_exact(e, lo, hi);
#pragma unroll
                            for (int i = 0; i < 3; i++) e.K[i] += dB[i];
                            sspan[tid][y] = lo <= hi ? (unsigned short)(lo | (hi << 8)) : (unsigned short)SPAN_EMPTY;
                            nrow += lo <= hi ? 1u : 0u;
                        }
#pragma unroll
                        for (int k = 0; k < 9; k++) szp[k][tid] = r.zp[k];
                        sid[tid] = id;
                    }
                }
                // phase 2: block prefix of the covered-row counts; every covered row becomes an item
                // (piece, row, first column, length), in piece-then-row order (neighbouring rows of
                // one piece have similar lengths, so a wave's items are of similar length)
                const u32 incl = wave_incl_scan(nrow);
                if ((tid & 63) == 63) sbucket[tid >> 6] = incl;
                __syncthreads();
                u32 pos0 = incl - nrow;
                for (int w = 0; w < (tid >> 6); w++) pos0 += sbucket[w];
                const u32 ni = sbucket[0] + sbucket[1] + sbucket[2] + sbucket[3];
                // the items go through LDS ITEM_CAP at a time (one round unless the batch is dense)
                for (u32 r0 = 0; r0 < ni; r0 += ITEM_CAP) {
                u32 pos = pos0;
                for (int y = yl; y <= yh; y++) {
                    u32 sp = sspan[tid][y];
                    int lo = (int)(sp & 0xffu), hi = (int)(sp >> 8);
                    if (lo > hi) continue;
                    if (pos >= r0 && pos < r0 + ITEM_CAP)
                        sitem[pos - r0] = (u32)tid | ((u32)y << 8) | ((u32)lo << 16) | ((u32)(hi - lo + 1) << 24);
                    pos++;
                }
                __syncthreads();
                // phase 3: thread = item
                const u32 nr = min((u32)ITEM_CAP, ni - r0);
                for (u32 i = tid; i < nr; i += 256) {
                    const u32 it = sitem[i];
                    const int j = (int)(it & 0xffu), y = (int)((it >> 8) & 0xffu), lo = (int)((it >> 16) & 0xffu), len = (int)(it >> 24);
                    const float pixy = spix[1][y];
                    const float z0 = szp[0][j], z2 = szp[2][j], w0 = szp[3][j], w2 = szp[5][j], d0 = szp[6][j], d2 = szp[8][j];
                    const float zy = szp[1][j] * pixy, wy = szp[4][j] * pixy, dy = szp[7][j] * pixy;
                    const u32 id = sid[j];
                    // branch-free body (the quotient is formed for every pixel, the minimum taken only
                    // where the tests pass); the pixel-centre x is recomputed (the same expression as
                    // spix) rather than read, so the loop holds no LDS read behind the LDS atomics
#pragma unroll 2
                    for (int x = lo; x < lo + len; x++) {
                        const float pixx = (float)((x0 + x + 0.5) * (double)a.xs - 1.0);
                        const float Zp = z0 * pixx + zy + z2;
                        const float Wp = w0 * pixx + wy + w2;
                        const float Dp = d0 * pixx + dy + d2;
                        const float z = Zp / Wp;
                        const bool ok = ((Wp > 0.0f) & (Dp > 0.0f)) | ((Wp < 0.0f) & (Dp < 0.0f));
                        if (ok & (z >= -1.0f) & (z <= 1.0f)) atomicMin(&key[y * TS + x], ((u64)ord_f32(z) << 32) | id);
                    }
                }
                __syncthreads();
                }
            }
            // the tile's winning keys: stored when this chunk is the tile's only one, else merged with
            // a device-scope 64-bit minimum (the key buffer starts at all-ones)
            const u32 begin = a.ws.offsets[tile], end = a.ws.offsets[tile + 1];
            const bool alone = scan_all || (end - begin) <= a.chunk;
            for (int i = tid; i < NP; i += 256) {
                const int px = x0 + i % TS, py = y0 + i / TS;
                if (px >= a.W || py >= a.H) continue;
                u64* dst = a.ws.keys + ((size_t)view * a.H + py) * a.W + px;
                const u64 k = key[i];
                if (alone) *dst = k;
                else if (k != ~0ull) atomicMin(dst, k);
            }
            __syncthreads();
        }
    }
};
template <int TS>
__global__ void __launch_bounds__(256, TS == 32 ? 4 : 6) k_fine(RastArgs a) { KFineF<TS>{}(vb_hw(), a); }

// One thread per pixel: winning key -> (u, v, z/w, id+1) and DB.
struct KResolveF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, RastArgs a) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t P = (size_t)a.W * a.H * a.depth;
        if (p >= P) return;
        const int px = (int)(p % a.W), py = (int)((p / a.W) % a.H);
        const u32 view = (u32)(p / ((size_t)a.W * a.H));
        fine_resolve(a, view, px, py, pixel_centre(px, a.xs), pixel_centre(py, a.ys), a.ws.keys[p]);
    }
};
__global__ void __launch_bounds__(256) k_resolve(RastArgs a) { KResolveF{}(vb_hw(), a); }

// RasterizeBackwardKernel (rasterize.cu:84-141).  One 16x16 pixel tile per workgroup; the
// nine (x, y, w) vertex terms of every covered pixel are summed per vertex in an LDS hash
// table, then each touched vertex gets one set of global atomics per tile.
constexpr int RB_SLOTS = 1024;

struct KRasterBwdF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int W, int H, int V, float xs, float ys, const float* proj, const u32* idx,
                                                    const float* rast, const float* grast, float* gproj, int tilesX) const {
        __shared__ LdsTable<RB_SLOTS, 3> tab;
        __shared__ u32 used[RB_SLOTS], wave_count[4];
        lds_table_clear(tab);
        __syncthreads();
        int view = vb.b.y;
        int tx = vb.b.x % tilesX, ty = vb.b.x / tilesX;
        int px = tx * 16 + (threadIdx.x & 15), py = ty * 16 + (threadIdx.x >> 4);
        const float4* clip = (const float4*)proj + (size_t)view * V;
        float* gp = gproj + (size_t)view * V * 4;
        bool act = false;
        int id = -1;
        u32 vi[3] = {0u, 0u, 0u};
        float gv[3][3];
        if (px < W && py < H) {
            size_t p = ((size_t)view * H + py) * W + px;
            float4 r = ((const float4*)rast)[p];
            id = (int)r.w - 1;
            float2 g = id >= 0 ? ((const float2*)grast)[p] : make_float2(0.f, 0.f);
            int bitcheck = __float_as_int(g.x) | __float_as_int(g.y);
            if (id >= 0 && (bitcheck << 1) != 0) {
                act = true;
                float pixx = (float)((px + 0.5) * (double)xs - 1.0);
                float pixy = (float)((py + 0.5) * (double)ys - 1.0);
                vi[0] = idx[3 * id]; vi[1] = idx[3 * id + 1]; vi[2] = idx[3 * id + 2];
                float4 c0 = clip[vi[0]], c1 = clip[vi[1]], c2 = clip[vi[2]];
                float p0x = c0.w * pixx - c0.x, p0y = c0.w * pixy - c0.y;
                float p1x = c1.w * pixx - c1.x, p1y = c1.w * pixy - c1.y;
                float p2x = c2.w * pixx - c2.x, p2y = c2.w * pixy - c2.y;
                float e0x = p0x - p2x, e0y = p0y - p2y, e1x = p1x - p2x, e1y = p1y - p2y;
                float aa = e0x * e1y - e0y * e1x;
                float ia = 1.f / (aa + (aa > 0 ? 1e-6f : -1e-6f));
                float kuv = r.x * g.x + r.y * g.y;
                gv[0][0] = (g.y * p2y + kuv * e1y) * ia;
                gv[0][1] = (-g.y * p2x - kuv * e1x) * ia;
                gv[0][2] = -gv[0][0] * pixx - gv[0][1] * pixy;
                gv[1][0] = (-g.x * p2y - kuv * e0y) * ia;
                gv[1][1] = (g.x * p2x + kuv * e0x) * ia;
                gv[1][2] = -gv[1][0] * pixx - gv[1][1] * pixy;
                gv[2][0] = ((-g.y * p0y + g.x * p1y) - kuv * (-p0y + p1y)) * ia;
                gv[2][1] = ((g.y * p0x - g.x * p1x) - kuv * (p0x - p1x)) * ia;
                gv[2][2] = -gv[2][0] * pixx - gv[2][1] * pixy;
            }
        }
        // Neighbouring pixels mostly share a triangle and so the same three vertex slots, whose LDS
        // atomics would serialise: the grads of a triangle covering RB_GROUP or more of the wave's
        // lanes are summed over the wave first and kept by its first lane, which alone adds them to
        // the table; lanes of smaller groups add their own.
        constexpr int RB_GROUP = 4;
        const int lane = threadIdx.x & 63;
        bool add = false;
        u64 pending = __ballot(act);
        while (pending) {
            int leader = __ffsll((unsigned long long)pending) - 1;
            int lid = __builtin_amdgcn_readlane(id, leader);
            bool mine = act && id == lid;
            u64 m = __ballot(mine);
            pending &= ~m;
            if (__popcll(m) >= RB_GROUP) {
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        float t = wave_sum_dpp(mine ? gv[k][j] : 0.f);
                        if (lane == leader) gv[k][j] = t;
                    }
                add |= lane == leader;
            } else {
                add |= mine;
            }
        }
        if (add) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                int sl = lds_table_slot(tab, vi[k]);
                if (sl >= 0) {
                    atomicAdd(&tab.val[0][sl], gv[k][0]);
                    atomicAdd(&tab.val[1][sl], gv[k][1]);
                    atomicAdd(&tab.val[2][sl], gv[k][2]);
                } else {
                    float* q = gp + 4 * (size_t)vi[k];
                    atomicAdd(q, gv[k][0]); atomicAdd(q + 1, gv[k][1]); atomicAdd(q + 3, gv[k][2]);
                }
            }
        }
        __syncthreads();
        lds_vertex_grad_flush(tab, used, wave_count, gp);
    }
};
__global__ void __launch_bounds__(256) k_raster_bwd(int W, int H, int V, float xs, float ys, const float* proj, const u32* idx,
                                                    const float* rast, const float* grast, float* gproj, int tilesX) { KRasterBwdF{}(vb_hw(), W, H, V, xs, ys, proj, idx, rast, grast, gproj, tilesX); }

}  // namespace

static int tile_shift_of(int tile) { return tile == 8 ? 3 : tile == 32 ? 5 : 4; }


// The workspace fits every tile size, so the tuning may change between calls.
extern "C" size_t vcl_rasterize_workspace_size(int width, int height, int depth, int projNum, int idxNum) {
    (void)projNum;
    if (width <= 0 || height <= 0 || depth <= 0 || idxNum < 0) return 0;
    size_t m = 0;
    for (int sh = 3; sh <= 5; sh++) m = std::max(m, make_layout(width, height, depth, idxNum, sh).total);
    return m;
}

int vcl_rasterize_forward_dev(const vcl_rasterize_params* p, int cleared, vcl_stream_t stream, const unsigned long long** keys) {
    VCL_REQUIRE(p && p->out && p->workspace && (p->idxNum == 0 || (p->proj && p->idx)));
    VCL_REQUIRE(p->width > 0 && p->height > 0 && p->depth > 0 && p->width <= 16384 && p->height <= 16384);
    VCL_REQUIRE(p->idxNum >= 0 && p->idxNum < (1 << 24) && (p->projNum > 0 || p->idxNum == 0) && p->depth < (1 << 23));
    VCL_REQUIRE(!p->enableDB || p->outDB);
    const int tshift = tile_shift_of(g_vcl_tuning.raster_tile);
    Layout L = make_layout(p->width, p->height, p->depth, p->idxNum, tshift);
    if (p->workspace_size < L.total) return VCL_ERR_OOM;
    hipStream_t s = as_hip(stream);
    RastArgs a;
    a.W = p->width; a.H = p->height; a.depth = p->depth; a.V = p->projNum; a.T = p->idxNum;
    a.enableDB = p->enableDB;
    a.xs = 2.f / (float)p->width; a.ys = 2.f / (float)p->height;
    a.proj = p->proj; a.idx = p->idx; a.out = p->out; a.outDB = p->enableDB ? p->outDB : nullptr;
    a.tilesX = L.tilesX; a.ntiles = L.ntiles; a.tshift = tshift; a.chunk = (u32)g_vcl_tuning.bin_chunk;
    a.nrec_direct = L.nrec_direct; a.ovf_cap = L.ovf_cap; a.bin_cap = L.bin_cap; a.chunk_cap = L.chunk_cap;
    a.ws = ws_ptrs(p->workspace, L);
    const size_t P = (size_t)a.W * a.H * a.depth;
    if (!cleared) {
        const WsClear c = ws_clear_of(a.ws, L, P);
        vlaunch<KClearF>(k_clear, dim3((unsigned)std::min<size_t>(cdiv(c.n(), 256), 4096)), dim3(256), 0, s, c);
        VCL_LAUNCH_CHECK();
    }
    if (L.nrec_direct > 0) {
        vlaunch<KSetupF>(k_setup, dim3(cdiv(L.nrec_direct, 256)), dim3(256), 0, s, a);
        VCL_LAUNCH_CHECK();
    }
    vlaunch<KBinLargeF<false>>(k_bin_large<false>, dim3(LARGE_BLOCKS), dim3(256), 0, s, a);
    VCL_LAUNCH_CHECK();
    vlaunch<KFoldLargeF>(k_fold_large, dim3(cdiv(L.NT, 256)), dim3(256), 0, s, a.ws, L.NT);
    VCL_LAUNCH_CHECK();
    vlaunch<KScanF>(k_scan, dim3(1), dim3(1024), 0, s, a, L.NT);
    VCL_LAUNCH_CHECK();
    {
        const u32 nfill = (u32)cdiv(L.nrec_cap, 256);
        vlaunch<KFillAllF>(k_fill_all, dim3(nfill + LARGE_BLOCKS), dim3(256), 0, s, a, L.nrec_cap, nfill);
        VCL_LAUNCH_CHECK();
    }
    const dim3 fg((unsigned)std::min(L.chunk_cap, (size_t)FINE_MAX_GRID));
    if (tshift == 3) vlaunch<KFineF<8>>(k_fine<8>, fg, dim3(256), 0, s, a);
    else if (tshift == 5) vlaunch<KFineF<32>>(k_fine<32>, fg, dim3(256), 0, s, a);
    else vlaunch<KFineF<16>>(k_fine<16>, fg, dim3(256), 0, s, a);
    VCL_LAUNCH_CHECK();
    if (keys) {
        *keys = (const unsigned long long*)a.ws.keys;
        return VCL_OK;
    }
    vlaunch<KResolveF>(k_resolve, dim3(cdiv(P, 256)), dim3(256), 0, s, a);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_rasterize_forward(const vcl_rasterize_params* p, vcl_stream_t stream) { return vcl_rasterize_forward_dev(p, 0, stream); }

int vcl_rasterize_clear_spec(const vcl_rasterize_params* p, WsClear* c) {
    VCL_REQUIRE(p && c && p->workspace && p->width > 0 && p->height > 0 && p->depth > 0 && p->idxNum >= 0);
    const int tshift = tile_shift_of(g_vcl_tuning.raster_tile);
    Layout L = make_layout(p->width, p->height, p->depth, p->idxNum, tshift);
    if (p->workspace_size < L.total) return VCL_ERR_OOM;
    *c = ws_clear_of(ws_ptrs(p->workspace, L), L, (size_t)p->width * p->height * p->depth);
    return VCL_OK;
}

extern "C" void vcl_rasterize_set_bin_capacity(size_t entries) { g_bin_cap_override = entries; }

extern "C" int vcl_rasterize_status(const vcl_rasterize_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->workspace);
    Layout L = make_layout(p->width, p->height, p->depth, p->idxNum, tile_shift_of(g_vcl_tuning.raster_tile));
    u32 misc[3];
    VCL_HIP_CHECK(hipMemcpyAsync(misc, (char*)p->workspace + L.off_misc, sizeof(misc), hipMemcpyDeviceToHost, as_hip(stream)));
    VCL_HIP_CHECK(hipStreamSynchronize(as_hip(stream)));
    return misc[2] ? VCL_ERR_OVERFLOW : VCL_OK;
}

extern "C" int vcl_rasterize_backward(const vcl_rasterize_params* p, const vcl_rasterize_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->proj && p->idx && p->out);
    if (!g->proj) return VCL_OK;
    VCL_REQUIRE(g->out);
    if ((size_t)p->width * p->height * p->depth == 0) return VCL_OK;
    int tilesX = (p->width + 15) / 16, tilesY = (p->height + 15) / 16;
    vlaunch<KRasterBwdF>(k_raster_bwd, dim3(tilesX * tilesY, p->depth), dim3(256), 0, as_hip(stream), p->width, p->height, p->projNum,
                       2.f / (float)p->width, 2.f / (float)p->height, p->proj, p->idx, p->out, g->out, g->proj, tilesX);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}
