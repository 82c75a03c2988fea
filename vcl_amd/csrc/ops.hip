// ops.hip -- per-pixel ops of the vcl hot path on gfx950: project, interpolate, mip texture,
// NormalAxis, ViewAxis, SGSpecular, SphericalGaussian, antialias, MSE, Adam and the mip
// utilities.  One thread per pixel over the row-major image (coalesced, background pixels exit
// early); reductions that the reference does with racy or atomic float adds are done with
// fixed-order block partials so results are deterministic.
#include "internal.hpp"
#include "raster_resolve.hpp"
#include <algorithm>
#include <cstring>
#include <cmath>
#include <mutex>

struct VclFillU32F {
    static constexpr int kBS = 1024, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, u32* p, size_t n, u32 v) const {
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t stride = (size_t)vb.g.x * blockDim.x;
        for (; i < n; i += stride) p[i] = v;
    }
};
__global__ void vcl_fill_u32(u32* p, size_t n, u32 v) { VclFillU32F{}(vb_hw(), p, n, v); }

int vcl_fill_u32_async(u32* p, size_t n, u32 v, hipStream_t s) {
    if (!p || n == 0) return VCL_OK;
    unsigned g = cdiv(n, 256);
    if (g > 4096) g = 4096;
    vlaunch<VclFillU32F>(vcl_fill_u32, dim3(g), dim3(256), 0, s, p, n, v);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

int vcl_fill_async(void* p, size_t bytes, hipStream_t s) { return vcl_fill_u32_async((u32*)p, bytes / 4, 0u, s); }

// Zero several buffers in one launch (each block clears 4096 words of one buffer).  Optionally
// block 0 also advances a pair of Adam bias-correction powers (bt[0] *= rhom, bt[1] *= rhov),
// so a step that starts with this clear needs no launch of its own for them.
constexpr int FILL_MAX = 8;
struct FillMulti {
    u32* p[FILL_MAX];
    size_t n[FILL_MAX];
    unsigned block_start[FILL_MAX + 1];
    int count;
    double* bt;
    double rhom, rhov;
};

struct VclFillMultiKF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, FillMulti f) const {
        unsigned b = vb.b.x;
        if (f.bt && b == 0 && threadIdx.x == 0) { f.bt[0] *= f.rhom; f.bt[1] *= f.rhov; }
        if (f.count == 0) return;
        int t = 0;
        while (t + 1 < f.count && b >= f.block_start[t + 1]) t++;
        size_t base = (size_t)(b - f.block_start[t]) * 4096;
        for (int k = threadIdx.x; k < 4096; k += 256) {
            size_t i = base + k;
            if (i < f.n[t]) f.p[t][i] = 0u;
        }
    }
};
__global__ void __launch_bounds__(256) vcl_fill_multi_k(FillMulti f) { VclFillMultiKF{}(vb_hw(), f); }

int vcl_fill_multi_async(void* const* p, const size_t* bytes, int n, hipStream_t s) {
    return vcl_fill_multi_bump_async(p, bytes, n, nullptr, 1.0, 1.0, s);
}

int vcl_fill_multi_bump_async(void* const* p, const size_t* bytes, int n, double* bt, double rhom, double rhov, hipStream_t s) {
    for (int base = 0; base < n || (base == 0 && bt); base += FILL_MAX) {
        FillMulti f;
        f.count = 0;
        f.bt = base == 0 ? bt : nullptr;
        f.rhom = rhom;
        f.rhov = rhov;
        unsigned blocks = 0;
        for (int i = base; i < n && i < base + FILL_MAX; i++) {
            if (!p[i] || bytes[i] < 4) continue;
            f.p[f.count] = (u32*)p[i];
            f.n[f.count] = bytes[i] / 4;
            f.block_start[f.count] = blocks;
            blocks += (unsigned)((f.n[f.count] + 4095) / 4096);
            f.count++;
        }
        f.block_start[f.count] = blocks;
        if (!blocks && !f.bt) continue;
        vlaunch<VclFillMultiKF>(vcl_fill_multi_k, dim3(blocks ? blocks : 1), dim3(256), 0, s, f);
        VCL_LAUNCH_CHECK();
    }
    return VCL_OK;
}

namespace {

constexpr int BS = 256;

// ---------------------------------------------------------------- deterministic column sums
// sum_r part[r * cols + c] in a fixed order: segment sums here, the sum over segments by the
// consumer (k_sg_lobe_final).
struct KColsumStageF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float* part, int rows, int cols, int seg_rows, float* seg) const {
        // block: 64 columns x 4 row groups; grid.y: segment
        int c = vb.b.x * 64 + (threadIdx.x & 63);
        int rg = threadIdx.x >> 6;
        int r0 = vb.b.y * seg_rows, r1 = min(rows, r0 + seg_rows);
        float s = 0.f;
        if (c < cols)
            for (int r = r0 + rg; r < r1; r += 4) s += part[(size_t)r * cols + c];
        __shared__ float sh[4][64];
        sh[rg][threadIdx.x & 63] = s;
        __syncthreads();
        if (rg == 0 && c < cols) seg[(size_t)vb.b.y * cols + c] = ((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) + sh[3][threadIdx.x];
    }
};
__global__ void __launch_bounds__(256) k_colsum_stage(const float* part, int rows, int cols, int seg_rows, float* seg) { KColsumStageF{}(vb_hw(), part, rows, cols, seg_rows, seg); }

constexpr int COLSUM_SEGS = 32;

size_t colsum_scratch_floats(int cols) { return (size_t)COLSUM_SEGS * cols; }

// First stage: segment sums of `rows` rows into seg_scratch.  colsum_stage_of describes it and
// returns the segment count; colsum_launch launches it (-1 when the launch failed).
int colsum_stage_of(const float* part, int rows, int cols, float* seg_scratch, VclColsumStage& cs) {
    int seg_rows = (rows + COLSUM_SEGS - 1) / COLSUM_SEGS;
    if (seg_rows < 1) seg_rows = 1;
    int nseg = (rows + seg_rows - 1) / seg_rows;
    if (nseg < 1) nseg = 1;
    cs.part = part; cs.rows = rows; cs.cols = cols; cs.seg_rows = seg_rows; cs.seg = seg_scratch;
    return nseg;
}
int colsum_nseg(const VclColsumStage& cs) { return (cs.rows + cs.seg_rows - 1) / cs.seg_rows; }
int colsum_launch(const VclColsumStage& cs, hipStream_t s) {
    vlaunch<KColsumStageF>(k_colsum_stage, dim3(cdiv(cs.cols, 64), colsum_nseg(cs)), dim3(256), 0, s, cs.part, cs.rows, cs.cols, cs.seg_rows, cs.seg);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}


// ---------------------------------------------------------------- project
// zero (optional): a buffer shaped like out cleared in the same pass (the step's clip-space grad
// clear); bt (optional): the step-start bump of Adam's bias-correction powers, done once.
struct KProjectFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int V, int dim, int views, const float* vec, const float* mat, float* out, float* zero,
                                                    double* bt, double rhom, double rhov, WsClear rc) const {
#pragma clang fp contract(off)
        size_t gid = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (bt && gid == 0) { bt[0] *= rhom; bt[1] *= rhov; }
        if (rc.counts && gid < rc.n()) ws_clear_elem(rc, gid);
        if (gid >= (size_t)V * views) return;
        int s = (int)(gid / V), v = (int)(gid % V);
        const float* m = mat + 16 * s;
        float x = vec[3 * v], y = vec[3 * v + 1], z = vec[3 * v + 2];
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = m[i] * x + m[4 + i] * y + m[8 + i] * z + m[12 + i];
        if (dim == 4) ((float4*)out)[gid] = make_float4(r[0], r[1], r[2], r[3]);
        else for (int i = 0; i < dim; i++) out[gid * dim + i] = r[i];
        if (zero) for (int i = 0; i < dim; i++) zero[gid * dim + i] = 0.f;
    }
};
__global__ void __launch_bounds__(BS) k_project_fwd(int V, int dim, int views, const float* vec, const float* mat, float* out, float* zero,
                                                    double* bt, double rhom, double rhov, WsClear rc) { KProjectFwdF{}(vb_hw(), V, dim, views, vec, mat, out, zero, bt, rhom, rhov, rc); }

// gvec[v] += sum over views of M^T g (views in order).  gmat (camera-matrix grad, off the
// fit path) is reduced per wave and added with one atomic per wave and entry.
struct KProjectBwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int V, int dim, int views, const float* vec, const float* mat, const float* gout,
                                                    float* gvec, float* gmat) const {
        int v = vb.b.x * blockDim.x + threadIdx.x;
        bool ok = v < V;
        float gx = 0, gy = 0, gz = 0;
        float pv[4] = {ok ? vec[3 * v] : 0.f, ok ? vec[3 * v + 1] : 0.f, ok ? vec[3 * v + 2] : 0.f, ok ? 1.f : 0.f};
        for (int s = 0; s < views; s++) {
            const float* m = mat + 16 * s;
            float g4[4];
            for (int i = 0; i < 4; i++) {
                float g = (ok && i < dim) ? gout[((size_t)s * V + v) * dim + i] : 0.f;
                g4[i] = g;
                gx += m[i] * g; gy += m[4 + i] * g; gz += m[8 + i] * g;
            }
            if (gmat) {
                for (int j = 0; j < 4; j++)
                    for (int i = 0; i < 4; i++) {
                        float val = wave_sum(pv[j] * g4[i]);
                        if ((threadIdx.x & 63) == 0 && val != 0.f) atomicAdd(gmat + 16 * s + 4 * j + i, val);
                    }
            }
        }
        if (ok && gvec) { gvec[3 * v] += gx; gvec[3 * v + 1] += gy; gvec[3 * v + 2] += gz; }
    }
};
__global__ void __launch_bounds__(BS) k_project_bwd(int V, int dim, int views, const float* vec, const float* mat, const float* gout,
                                                    float* gvec, float* gmat) { KProjectBwdF{}(vb_hw(), V, dim, views, vec, mat, gout, gvec, gmat); }

// ---------------------------------------------------------------- interpolate
// One attribute of one covered pixel (interpolate.cu:38: a0*u + a1*v in float, the (1.0 - u - v)
// term w2 in double) and its screen derivatives from the barycentric derivatives b.
__device__ __forceinline__ void interp_attr_px(float4 r, float4 b, double w2, float a0, float a1, float a2, float& o, float2& oda) {
#pragma clang fp contract(off)
    o = (float)((double)(a0 * r.x + a1 * r.y) + (double)a2 * w2);
    float dadu = a0 - a2, dadv = a1 - a2;
    oda = make_float2(dadu * b.x + dadv * b.z, dadu * b.y + dadv * b.w);
}

struct KInterpFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int D, int enableDA, const float* rast, const float* db, const float* attr,
                                                   const u32* idx, float* out, float* outDA) const {
#pragma clang fp contract(off)
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float4 r = ((const float4*)rast)[p];
        int id = (int)r.w - 1;
        if (id < 0) {
            for (int i = 0; i < D; i++) {
                out[p * D + i] = 0.f;
                if (enableDA) ((float2*)outDA)[p * D + i] = make_float2(0.f, 0.f);
            }
            return;
        }
        u32 i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
        float4 b = enableDA ? ((const float4*)db)[p] : make_float4(0, 0, 0, 0);
        double w2 = 1.0 - (double)r.x - (double)r.y;
        for (int i = 0; i < D; i++) {
            float o;
            float2 oda;
            interp_attr_px(r, b, w2, attr[(size_t)i0 * D + i], attr[(size_t)i1 * D + i], attr[(size_t)i2 * D + i], o, oda);
            out[p * D + i] = o;
            if (enableDA) ((float2*)outDA)[p * D + i] = oda;
        }
    }
};
__global__ void __launch_bounds__(BS) k_interp_fwd(size_t P, int D, int enableDA, const float* rast, const float* db, const float* attr,
                                                   const u32* idx, float* out, float* outDA) { KInterpFwdF{}(vb_hw(), P, D, enableDA, rast, db, attr, idx, out, outDA); }

struct KInterpBwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int D, const float* rast, const float* attr, const u32* idx, const float* gout,
                                                   float* gattr, float* grast) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float4 r = ((const float4*)rast)[p];
        int id = (int)r.w - 1;
        if (id < 0) return;
        u32 i0 = idx[3 * id], i1 = idx[3 * id + 1], i2 = idx[3 * id + 2];
        float gu = 0.f, gv = 0.f;
        float w2 = 1.0f - r.x - r.y;
        for (int i = 0; i < D; i++) {
            float g = gout[p * D + i];
            if (gattr) {
                atomicAdd(gattr + (size_t)i0 * D + i, g * r.x);
                atomicAdd(gattr + (size_t)i1 * D + i, g * r.y);
                atomicAdd(gattr + (size_t)i2 * D + i, g * w2);
            }
            float a2 = attr[(size_t)i2 * D + i];
            gu += g * (attr[(size_t)i0 * D + i] - a2);
            gv += g * (attr[(size_t)i1 * D + i] - a2);
        }
        if (grast) ((float2*)grast)[p] = make_float2(gu, gv);
    }
};
__global__ void __launch_bounds__(BS) k_interp_bwd(size_t P, int D, const float* rast, const float* attr, const u32* idx, const float* gout,
                                                   float* gattr, float* grast) { KInterpBwdF{}(vb_hw(), P, D, rast, attr, idx, gout, gattr, grast); }

// ---------------------------------------------------------------- mip texture
struct TexArgs {
    int tw, th, C, L;
    const float* tex[VCL_TEX_MAX_MIP_LEVEL];
};
struct TexGradArgs {
    float* tex[VCL_TEX_MAX_MIP_LEVEL];
};

__device__ __forceinline__ void calc_level(float4 da, int tw, int th, int L, int& l0, int& l1, float& f, float& level, float4* dl, int mode) {
    float dsdx = da.x * tw, dsdy = da.y * tw, dtdx = da.z * th, dtdy = da.w * th;
    float s2 = dsdx * dsdx + dsdy * dsdy, t2 = dtdx * dtdx + dtdy * dtdy;
    float b = .5f * (s2 + t2);
    // c^2 = b^2 - a^2 = (b - a)(b + a), each factor a sum of squares (no cancellation)
    float pm = dsdx - dtdy, qm = dsdy + dtdx, pp = dsdx + dtdy, qp = dsdy - dtdx;
    float c2 = .25f * ((pm * pm + qm * qm) * (pp * pp + qp * qp));
    float c = sqrtf(c2);
    level = .5f * log2f(b + c);
    l0 = level <= 0 ? 0 : (float)(L - 2) <= level ? L - 2 : (int)floorf(level);
    l1 = level <= 1 ? 1 : (float)(L - 1) <= level ? L - 1 : (int)floorf(level) + 1;
    f = level <= 0 ? 0.f : (float)(L - 1) <= level ? 1.f : level - floorf(level);
    if (!dl) return;
    float a = dsdx * dtdy - dtdx * dsdy;
    if (mode == 0) {
        float d = b * c + c2;
        if (fabsf(d) > 1e-6) {
            d = 0.72134752f / d;
            float bc = b + c;
            *dl = make_float4(d * (bc * dsdx - a * dtdy), d * (bc * dsdy + a * dtdx), (bc * dtdx + a * dsdy), (bc * dtdy - a * dsdx));
        } else if (fabsf(b) > 1e-6) {
            d = 1.f / b;
            *dl = make_float4(d * dsdx, d * dsdy, d * dtdx, d * dtdy);
        } else {
            *dl = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    } else {
        float k = 1.f / (2.f * 0.69314718055994530942f * (b + c));
        bool live = level > 0 && level < (float)(L - 1) && c > 0.f && isfinite(k);
        if (live) {
            float ic = 1.f / c;
            float p0 = (dsdx + (b * dsdx - a * dtdy) * ic) * k;
            float p1 = (dsdy + (b * dsdy + a * dtdx) * ic) * k;
            float p2 = (dtdx + (b * dtdx + a * dsdy) * ic) * k;
            float p3 = (dtdy + (b * dtdy - a * dsdx) * ic) * k;
            *dl = make_float4(p0 * tw, p1 * tw, p2 * th, p3 * th);
        } else {
            *dl = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

__device__ __forceinline__ void index_fetch(int tw, int th, int level, float2 uv, int4& id, float& fx, float& fy) {
    int sx = tw >> level, sy = th >> level;
    float tx = uv.x * (float)sx, ty = uv.y * (float)sy;
    int u0 = tx < 0 ? 0 : tx > sx - 1 ? sx - 1 : (int)tx;
    int u1 = tx < 1 ? 0 : tx > sx - 2 ? sx - 1 : (int)tx + 1;
    int v0 = ty < 0 ? 0 : ty > sy - 1 ? sy - 1 : (int)ty;
    int v1 = ty < 1 ? 0 : ty > sy - 2 ? sy - 1 : (int)ty + 1;
    id = make_int4(v0 * sx + u0, v0 * sx + u1, v1 * sx + u0, v1 * sx + u1);
    fx = tx < 0 ? 0.f : (float)sx < tx ? 1.f : tx - floorf(tx);
    fy = ty < 0 ? 0.f : (float)sy < ty ? 1.f : ty - floorf(ty);
}

__device__ __forceinline__ float lerpf_(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(lerpf_(a.x, b.x, t), lerpf_(a.y, b.y, t), lerpf_(a.z, b.z, t), lerpf_(a.w, b.w, t));
}
__device__ __forceinline__ float4 bilerp4(float4 a, float4 b, float4 c, float4 d, float tx, float ty) {
    return lerp4(lerp4(a, b, tx), lerp4(c, d, tx), ty);
}

// C == 4 fast path: texels as float4.
// C == 4 texture sample of one covered pixel at uv u with screen derivatives da (used when
// t.L > 1): trilinear between the two mip levels around the footprint's level.
__device__ __forceinline__ float4 tex4_px(const TexArgs& t, float2 u, float4 da) {
    float4 o;
    {
        if (t.L > 1) {
            int l0, l1; float f, lev;
            calc_level(da, t.tw, t.th, t.L, l0, l1, f, lev, nullptr, 0);
            int4 i0; float fx, fy;
            index_fetch(t.tw, t.th, l0, u, i0, fx, fy);
            const float4* T0 = (const float4*)t.tex[l0];
            o = bilerp4(T0[i0.x], T0[i0.y], T0[i0.z], T0[i0.w], fx, fy);
            if (f > 0) {
                int4 i1; float gx, gy;
                index_fetch(t.tw, t.th, l1, u, i1, gx, gy);
                const float4* T1 = (const float4*)t.tex[l1];
                o = lerp4(o, bilerp4(T1[i1.x], T1[i1.y], T1[i1.z], T1[i1.w], gx, gy), f);
            }
        } else {
            int4 i0; float fx, fy;
            index_fetch(t.tw, t.th, 0, u, i0, fx, fy);
            const float4* T0 = (const float4*)t.tex[0];
            o = bilerp4(T0[i0.x], T0[i0.y], T0[i0.z], T0[i0.w], fx, fy);
        }
    }
    return o;
}

struct KTexFwd4F {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, TexArgs t, const float* rast, const float* uv, const float* uvDA, float* out) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
        if (rast[4 * p + 3] >= 1.f)
            o = tex4_px(t, ((const float2*)uv)[p], t.L > 1 ? ((const float4*)uvDA)[p] : make_float4(0.f, 0.f, 0.f, 0.f));
        ((float4*)out)[p] = o;
    }
};
__global__ void __launch_bounds__(BS) k_tex_fwd4(size_t P, TexArgs t, const float* rast, const float* uv, const float* uvDA, float* out) { KTexFwd4F{}(vb_hw(), P, t, rast, uv, uvDA, out); }

// Texel addresses of one C == 4 sample (tex4_px split so that the loads of several pixels can be
// issued before any is used): level l0 at i0 with weights fx, fy; when two, also level l1 at i1
// (gx, gy) blended by f.
struct Tex4Taps {
    const float4* T0;
    const float4* T1;
    int4 i0, i1;
    float fx, fy, gx, gy, f;
    bool two;
};

__device__ __forceinline__ Tex4Taps tex4_taps(const TexArgs& t, float2 u, float4 da) {
    Tex4Taps q;
    q.two = false;
    q.f = 0.f;
    q.gx = q.gy = 0.f;
    if (t.L > 1) {
        int l0, l1; float lev;
        calc_level(da, t.tw, t.th, t.L, l0, l1, q.f, lev, nullptr, 0);
        index_fetch(t.tw, t.th, l0, u, q.i0, q.fx, q.fy);
        q.T0 = (const float4*)t.tex[l0];
        q.two = q.f > 0;
        index_fetch(t.tw, t.th, l1, u, q.i1, q.gx, q.gy);
        q.T1 = (const float4*)t.tex[l1];
    } else {
        index_fetch(t.tw, t.th, 0, u, q.i0, q.fx, q.fy);
        q.T0 = (const float4*)t.tex[0];
        q.T1 = q.T0;
        q.i1 = q.i0;
    }
    return q;
}

// The fit step's uv interpolation (2 attributes, with derivatives) and C == 4 texture forward in
// one pass: the same per-pixel arithmetic as the two ops (interp_attr_px, tex4_px), writing uv and
// its derivatives for the backward and the texture sample.  Each thread takes NPX pixels one
// workgroup-width apart and goes through the dependent loads in stages (rast; triangle indices
// and DB; uv attributes; texels), each stage's loads for all its pixels issued together, so a
// wave keeps NPX chains in flight instead of one.
// zero (optional): a buffer shaped like out (4 floats per pixel) cleared in the same pass.
template <int NPX>
struct KInterpTex4F {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, const float* rast, const float* db, const float* uvattr, const u32* idx, float* uvp,
                                               float* uvDA, TexArgs t, float* out, float* zero) const {
        const size_t p0 = (size_t)vb.b.x * (BS * NPX) + threadIdx.x;
        float4 r[NPX];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            r[j] = p < P ? ((const float4*)rast)[p] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        u32 i0[NPX], i1[NPX], i2[NPX];
        float4 b[NPX];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const int id = (int)r[j].w - 1;
            i0[j] = i1[j] = i2[j] = 0;
            b[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (id >= 0) {
                i0[j] = idx[3 * id]; i1[j] = idx[3 * id + 1]; i2[j] = idx[3 * id + 2];
                b[j] = ((const float4*)db)[p0 + (size_t)j * BS];
            }
        }
        float2 a0[NPX], a1[NPX], a2[NPX];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            a0[j] = a1[j] = a2[j] = make_float2(0.f, 0.f);
            if ((int)r[j].w - 1 >= 0) {
                a0[j] = ((const float2*)uvattr)[i0[j]];
                a1[j] = ((const float2*)uvattr)[i1[j]];
                a2[j] = ((const float2*)uvattr)[i2[j]];
            }
        }
        Tex4Taps q[NPX];
        float4 tx[NPX][8];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            float2 u = make_float2(0.f, 0.f);
            float4 da = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((int)r[j].w - 1 >= 0) {
                const double w2 = 1.0 - (double)r[j].x - (double)r[j].y;
                float2 dau, dav;
                interp_attr_px(r[j], b[j], w2, a0[j].x, a1[j].x, a2[j].x, u.x, dau);
                interp_attr_px(r[j], b[j], w2, a0[j].y, a1[j].y, a2[j].y, u.y, dav);
                da = make_float4(dau.x, dau.y, dav.x, dav.y);
            }
            if (p < P) {
                ((float2*)uvp)[p] = u;
                ((float4*)uvDA)[p] = da;
            }
            if (r[j].w >= 1.f) {
                q[j] = tex4_taps(t, u, da);
                tx[j][0] = q[j].T0[q[j].i0.x]; tx[j][1] = q[j].T0[q[j].i0.y]; tx[j][2] = q[j].T0[q[j].i0.z]; tx[j][3] = q[j].T0[q[j].i0.w];
                if (q[j].two) {
                    tx[j][4] = q[j].T1[q[j].i1.x]; tx[j][5] = q[j].T1[q[j].i1.y]; tx[j][6] = q[j].T1[q[j].i1.z]; tx[j][7] = q[j].T1[q[j].i1.w];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            if (p >= P) continue;
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r[j].w >= 1.f) {
                o = bilerp4(tx[j][0], tx[j][1], tx[j][2], tx[j][3], q[j].fx, q[j].fy);
                if (q[j].two) o = lerp4(o, bilerp4(tx[j][4], tx[j][5], tx[j][6], tx[j][7], q[j].gx, q[j].gy), q[j].f);
            }
            ((float4*)out)[p] = o;
            if (zero) ((float4*)zero)[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
};
constexpr int INTERP_TEX_NPX = 2;
__global__ void __launch_bounds__(BS) k_interp_tex4(size_t P, const float* rast, const float* db, const float* uvattr, const u32* idx, float* uvp,
                                                    float* uvDA, TexArgs t, float* out, float* zero) {
    KInterpTex4F<INTERP_TEX_NPX>{}(vb_hw(), P, rast, db, uvattr, idx, uvp, uvDA, t, out, zero);
}

struct KTexFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, TexArgs t, const float* rast, const float* uv, const float* uvDA, float* out) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        int C = t.C;
        if (rast[4 * p + 3] < 1.f) { for (int c = 0; c < C; c++) out[p * C + c] = 0.f; return; }
        float2 u = ((const float2*)uv)[p];
        int l0 = 0, l1 = 0; float f = 0.f, lev;
        if (t.L > 1) calc_level(((const float4*)uvDA)[p], t.tw, t.th, t.L, l0, l1, f, lev, nullptr, 0);
        int4 i0, i1; float fx, fy, gx = 0, gy = 0;
        index_fetch(t.tw, t.th, l0, u, i0, fx, fy);
        if (t.L > 1) index_fetch(t.tw, t.th, l1, u, i1, gx, gy);
        const float* T0 = t.tex[l0];
        const float* T1 = t.tex[l1];
        for (int c = 0; c < C; c++) {
            float v = lerpf_(lerpf_(T0[i0.x * C + c], T0[i0.y * C + c], fx), lerpf_(T0[i0.z * C + c], T0[i0.w * C + c], fx), fy);
            if (t.L > 1 && f > 0) {
                float v1 = lerpf_(lerpf_(T1[i1.x * C + c], T1[i1.y * C + c], gx), lerpf_(T1[i1.z * C + c], T1[i1.w * C + c], gx), gy);
                v = lerpf_(v, v1, f);
            }
            out[p * C + c] = v;
        }
    }
};
__global__ void __launch_bounds__(BS) k_tex_fwd(size_t P, TexArgs t, const float* rast, const float* uv, const float* uvDA, float* out) { KTexFwdF{}(vb_hw(), P, t, rast, uv, uvDA, out); }

struct KTexBwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, TexArgs t, TexGradArgs g, const float* rast, const float* uv, const float* uvDA,
                                                const float* gout, float* guv, float* guvDA, int mode) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        if (rast[4 * p + 3] < 1.f) return;
        int C = t.C;
        float2 u = ((const float2*)uv)[p];
        if (t.L <= 1) {
            int4 i0; float fx, fy;
            index_fetch(t.tw, t.th, 0, u, i0, fx, fy);
            float gu = 0.f, gv = 0.f;
            const float* T0 = t.tex[0];
            float* G0 = g.tex[0];
            for (int c = 0; c < C; c++) {
                float go = gout[p * C + c];
                if (G0) {
                    atomicAdd(G0 + i0.x * C + c, (1.f - fx) * (1.f - fy) * go);
                    atomicAdd(G0 + i0.y * C + c, fx * (1.f - fy) * go);
                    atomicAdd(G0 + i0.z * C + c, (1.f - fx) * fy * go);
                    atomicAdd(G0 + i0.w * C + c, fx * fy * go);
                }
                float t00 = T0[i0.x * C + c], t01 = T0[i0.y * C + c], t10 = T0[i0.z * C + c], t11 = T0[i0.w * C + c];
                gu += lerpf_(t01 - t00, t11 - t10, fy) * (float)t.tw * go;
                gv += lerpf_(t10 - t00, t11 - t01, fx) * (float)t.th * go;
            }
            if (guv) ((float2*)guv)[p] = make_float2(gu, gv);
            return;
        }
        int l0, l1; float f, lev; float4 dl;
        calc_level(((const float4*)uvDA)[p], t.tw, t.th, t.L, l0, l1, f, lev, &dl, mode);
        int4 i0, i1; float fx, fy, hx, hy;
        index_fetch(t.tw, t.th, l0, u, i0, fx, fy);
        index_fetch(t.tw, t.th, l1, u, i1, hx, hy);
        const float* T0 = t.tex[l0];
        const float* T1 = t.tex[l1];
        float* G0 = g.tex[l0];
        float* G1 = g.tex[l1];
        float gu = 0.f, gv = 0.f, gl = 0.f;
        float w0 = 1.f - f;
        for (int c = 0; c < C; c++) {
            float go = gout[p * C + c];
            if (G0) {
                atomicAdd(G0 + i0.x * C + c, w0 * (1.f - fx) * (1.f - fy) * go);
                atomicAdd(G0 + i0.y * C + c, w0 * fx * (1.f - fy) * go);
                atomicAdd(G0 + i0.z * C + c, w0 * (1.f - fx) * fy * go);
                atomicAdd(G0 + i0.w * C + c, w0 * fx * fy * go);
            }
            float t00 = T0[i0.x * C + c], t01 = T0[i0.y * C + c], t10 = T0[i0.z * C + c], t11 = T0[i0.w * C + c];
            float du = lerpf_(t01 - t00, t11 - t10, fy) * (float)(t.tw >> l0);
            float dv = lerpf_(t10 - t00, t11 - t01, fx) * (float)(t.th >> l0);
            if (f > 0) {
                float lv = lerpf_(lerpf_(t00, t01, fx), lerpf_(t10, t11, fx), fy);
                if (G1) {
                    atomicAdd(G1 + i1.x * C + c, f * (1.f - hx) * (1.f - hy) * go);
                    atomicAdd(G1 + i1.y * C + c, f * hx * (1.f - hy) * go);
                    atomicAdd(G1 + i1.z * C + c, f * (1.f - hx) * hy * go);
                    atomicAdd(G1 + i1.w * C + c, f * hx * hy * go);
                }
                float s00 = T1[i1.x * C + c], s01 = T1[i1.y * C + c], s10 = T1[i1.z * C + c], s11 = T1[i1.w * C + c];
                du = lerpf_(du, lerpf_(s01 - s00, s11 - s10, hy) * (float)(t.tw >> l1), f);
                dv = lerpf_(dv, lerpf_(s10 - s00, s11 - s01, hx) * (float)(t.th >> l1), f);
                gl += (lerpf_(lerpf_(s00, s01, hx), lerpf_(s10, s11, hx), hy) - lv) * go;
            }
            gu += du * go;
            gv += dv * go;
        }
        if (guv) ((float2*)guv)[p] = make_float2(gu, gv);
        if (guvDA) ((float4*)guvDA)[p] = make_float4(gl * dl.x, gl * dl.y, gl * dl.z, gl * dl.w);
    }
};
__global__ void __launch_bounds__(BS) k_tex_bwd(size_t P, TexArgs t, TexGradArgs g, const float* rast, const float* uv, const float* uvDA,
                                                const float* gout, float* guv, float* guvDA, int mode) { KTexBwdF{}(vb_hw(), P, t, g, rast, uv, uvDA, gout, guv, guvDA, mode); }

// C == 4: texel grads of a workgroup's pixels are summed in an LDS table of NB buckets, one
// bucket per run of 4 consecutive texels of a level (one 64-byte line of float4 grads), keyed by
// (level, texel / 4).  The flush compacts the occupied buckets and hands each to 16 consecutive
// lanes, texel-major and channel-minor, so one wave's atomics cover 4 whole lines instead of 64
// scattered dwords.  A key that finds no bucket within the probe limit falls back to direct
// global atomics.
//
// The table sums in 64-bit fixed point (fixed64), scaled by M, the largest |go| component over
// the workgroup's covered pixels: every contribution w * go[c] is at most M in size, a pixel adds
// less than 2^52 + 4 to one texel channel (its bilinear weights at one level sum to at most 1),
// and a sum over the at most 1024 pixels of a workgroup stays below 2^63.  Each contribution is
// truncated by less than 2^-51 M (float atomics round every partial sum to 2^-24 of its own
// size).  A workgroup whose go holds a non-finite value sends its splats to global float atomics.
constexpr u32 TB_LEVEL_SHIFT = 22;  // key = level << 22 | texel >> 2 (texel < 2^24)

template <int NB>
struct TexGradTable {
    unsigned long long val[NB * 16];  // [bucket][texel & 3][channel], fixed point (2^-S units)
    u32 key[NB];
    u32 used[NB];        // occupied buckets, compacted before the flush
    u32 wave_count[BS / 64];
    u32 wave_max[BS / 64];  // per wave: the largest |go| component, as its bit pattern
};

template <int NB>
__device__ __forceinline__ void tex_table_clear(TexGradTable<NB>& t) {
    for (int i = threadIdx.x; i < NB; i += blockDim.x) t.key[i] = LDS_EMPTY;
    ulonglong2* v = (ulonglong2*)t.val;
    for (int i = threadIdx.x; i < NB * 8; i += blockDim.x) v[i] = make_ulonglong2(0ull, 0ull);
}

// Keys only go from empty to their final value, so a plain read that finds k is final; the CAS
// is issued only for an empty bucket (a read of a stale "empty" is settled by the CAS).
template <int NB>
__device__ __forceinline__ int tex_table_bucket(TexGradTable<NB>& t, u32 k) {
    u32 h = fib_hash<ilog2_c(NB)>(k);
    for (int probe = 0; probe < 24; probe++) {
        u32 cur = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == k) return (int)h;
        if (cur == LDS_EMPTY) {
            u32 old = atomicCAS(&t.key[h], LDS_EMPTY, k);
            if (old == LDS_EMPTY || old == k) return (int)h;
        }
        h = (h + 1) & (NB - 1);
    }
    return -1;
}

// One texel's splat of weight w: into the table (gs = go * 2^S), or, for direct, by global float
// atomics of w * go.
template <int NB>
__device__ __forceinline__ void tex_splat(TexGradTable<NB>& tab, float* G, int level, int texel, float w, float4 go, float4 gs, bool direct) {
    const int b = direct ? -1 : tex_table_bucket(tab, ((u32)level << TB_LEVEL_SHIFT) | ((u32)texel >> 2));
    if (b >= 0) {
        unsigned long long* q = tab.val + b * 16 + (texel & 3) * 4;
        atomicAdd(q, fixed64(w * gs.x));
        atomicAdd(q + 1, fixed64(w * gs.y));
        atomicAdd(q + 2, fixed64(w * gs.z));
        atomicAdd(q + 3, fixed64(w * gs.w));
    } else {
        float* q = G + (size_t)texel * 4;
        atomicAdd(q, w * go.x); atomicAdd(q + 1, w * go.y); atomicAdd(q + 2, w * go.z); atomicAdd(q + 3, w * go.w);
    }
}

// Flush of the table (after a barrier): the occupied buckets are compacted into used[], then
// lane j of a bucket's 16 adds dword j of its 64-byte run (texel j / 4, channel j % 4), converted
// back from fixed point.  Untouched texels of a run hold exact zeros and are skipped, which also
// keeps runs that pass the end of a small level (fewer than 4 texels) inside it.
template <int NB>
__device__ __forceinline__ void tex_table_flush(TexGradTable<NB>& tab, const TexGradArgs& g, int S) {
    u32 n = 0;
    for (int c = 0; c < NB; c += BS) {
        int b = c + threadIdx.x;
        block_append(b < NB && tab.key[b] != LDS_EMPTY, (u32)b, tab.used, tab.wave_count, n);
    }
    for (u32 i = threadIdx.x; i < n * 16; i += BS) {
        u32 b = tab.used[i >> 4];
        const long long q = (long long)tab.val[b * 16 + (i & 15)];
        if (q == 0) continue;
        u32 k = tab.key[b];
        size_t dword = ((size_t)(k & ((1u << TB_LEVEL_SHIFT) - 1)) << 4) | (i & 15);
        atomicAdd(g.tex[k >> TB_LEVEL_SHIFT] + dword, ldexpf((float)q, -S));
    }
}

// The interpolate backward of a 2-channel attribute fused into the texture backward that produces
// its output grad (grast null: not fused).
struct InterpBwdArgs {
    const float* rast;
    const u32* idx;
    const float* attr;
    float* grast;
};

// One covered pixel p of the C == 4 texture backward, with its output grad go (gs: go in the
// table's fixed-point units): texel splats (want_tex), uv / uvDA grads stored, and, when fused,
// the interpolate backward's rast grad from the uv grad.
template <int NB>
__device__ __forceinline__ void tex_bwd4_pixel(TexGradTable<NB>& tab, bool want_tex, bool direct, size_t p, float4 go, float4 gs,
                                               const TexArgs& t, const TexGradArgs& g, const float* uv, const float* uvDA, float* guv,
                                               float* guvDA, int mode, const InterpBwdArgs& ib) {
    // the fused interpolate backward's triangle attributes, loaded first so that their chain of
    // loads overlaps the texture work
    float a0x = 0.f, a0y = 0.f, a1x = 0.f, a1y = 0.f, a2x = 0.f, a2y = 0.f;
    if (ib.grast) {
        const int id = (int)ib.rast[4 * p + 3] - 1;
        const u32 i0 = ib.idx[3 * id], i1 = ib.idx[3 * id + 1], i2 = ib.idx[3 * id + 2];
        a0x = ib.attr[2 * (size_t)i0]; a0y = ib.attr[2 * (size_t)i0 + 1];
        a1x = ib.attr[2 * (size_t)i1]; a1y = ib.attr[2 * (size_t)i1 + 1];
        a2x = ib.attr[2 * (size_t)i2]; a2y = ib.attr[2 * (size_t)i2 + 1];
    }
    float2 u = ((const float2*)uv)[p];
    int l0 = 0, l1 = 0; float f = 0.f, lev; float4 dl = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t.L > 1) calc_level(((const float4*)uvDA)[p], t.tw, t.th, t.L, l0, l1, f, lev, guvDA ? &dl : nullptr, mode);
    int4 i0; float fx, fy;
    index_fetch(t.tw, t.th, l0, u, i0, fx, fy);
    const float4* T0 = (const float4*)t.tex[l0];
    float4 a00 = T0[i0.x], a01 = T0[i0.y], a10 = T0[i0.z], a11 = T0[i0.w];
    float w0 = 1.f - f;
    if (want_tex) {
        float* G0 = g.tex[l0];
        tex_splat(tab, G0, l0, i0.x, w0 * (1.f - fx) * (1.f - fy), go, gs, direct);
        tex_splat(tab, G0, l0, i0.y, w0 * fx * (1.f - fy), go, gs, direct);
        tex_splat(tab, G0, l0, i0.z, w0 * (1.f - fx) * fy, go, gs, direct);
        tex_splat(tab, G0, l0, i0.w, w0 * fx * fy, go, gs, direct);
    }
    float s0 = (float)(t.tw >> l0), s1 = (float)(t.th >> l0);
    float gu = (lerpf_(a01.x - a00.x, a11.x - a10.x, fy) * s0) * go.x + (lerpf_(a01.y - a00.y, a11.y - a10.y, fy) * s0) * go.y +
               (lerpf_(a01.z - a00.z, a11.z - a10.z, fy) * s0) * go.z + (lerpf_(a01.w - a00.w, a11.w - a10.w, fy) * s0) * go.w;
    float gv = (lerpf_(a10.x - a00.x, a11.x - a01.x, fx) * s1) * go.x + (lerpf_(a10.y - a00.y, a11.y - a01.y, fx) * s1) * go.y +
               (lerpf_(a10.z - a00.z, a11.z - a01.z, fx) * s1) * go.z + (lerpf_(a10.w - a00.w, a11.w - a01.w, fx) * s1) * go.w;
    float gl = 0.f;
    if (t.L > 1 && f > 0) {
        int4 i1; float hx, hy;
        index_fetch(t.tw, t.th, l1, u, i1, hx, hy);
        const float4* T1 = (const float4*)t.tex[l1];
        float4 b00 = T1[i1.x], b01 = T1[i1.y], b10 = T1[i1.z], b11 = T1[i1.w];
        if (want_tex) {
            float* G1 = g.tex[l1];
            tex_splat(tab, G1, l1, i1.x, f * (1.f - hx) * (1.f - hy), go, gs, direct);
            tex_splat(tab, G1, l1, i1.y, f * hx * (1.f - hy), go, gs, direct);
            tex_splat(tab, G1, l1, i1.z, f * (1.f - hx) * hy, go, gs, direct);
            tex_splat(tab, G1, l1, i1.w, f * hx * hy, go, gs, direct);
        }
        float r0 = (float)(t.tw >> l1), r1 = (float)(t.th >> l1);
        // per channel: lerp(level-0 derivative, level-1 derivative, f), as texturemap.cu:273-274
        float du[4], dv[4];
        float ax[4] = {a00.x, a00.y, a00.z, a00.w}, bx[4] = {a01.x, a01.y, a01.z, a01.w};
        float cx[4] = {a10.x, a10.y, a10.z, a10.w}, dx[4] = {a11.x, a11.y, a11.z, a11.w};
        float ex[4] = {b00.x, b00.y, b00.z, b00.w}, fx1[4] = {b01.x, b01.y, b01.z, b01.w};
        float gx1[4] = {b10.x, b10.y, b10.z, b10.w}, hx1[4] = {b11.x, b11.y, b11.z, b11.w};
        float gg[4] = {go.x, go.y, go.z, go.w};
        gu = 0.f; gv = 0.f;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            du[c] = lerpf_(lerpf_(bx[c] - ax[c], dx[c] - cx[c], fy) * s0, lerpf_(fx1[c] - ex[c], hx1[c] - gx1[c], hy) * r0, f);
            dv[c] = lerpf_(lerpf_(cx[c] - ax[c], dx[c] - bx[c], fx) * s1, lerpf_(gx1[c] - ex[c], hx1[c] - fx1[c], hx) * r1, f);
            float lv = lerpf_(lerpf_(ax[c], bx[c], fx), lerpf_(cx[c], dx[c], fx), fy);
            gl += (lerpf_(lerpf_(ex[c], fx1[c], hx), lerpf_(gx1[c], hx1[c], hx), hy) - lv) * gg[c];
            gu += du[c] * gg[c];
            gv += dv[c] * gg[c];
        }
    }
    if (guv) ((float2*)guv)[p] = make_float2(gu, gv);
    if (guvDA && t.L > 1) ((float4*)guvDA)[p] = make_float4(gl * dl.x, gl * dl.y, gl * dl.z, gl * dl.w);
    if (ib.grast) {
        // InterpolateBackward (interpolate.cu:114-143) for D = 2 with output grad (gu, gv), in
        // its channel order
        float ru = 0.f, rv = 0.f;
        ru += gu * (a0x - a2x); rv += gu * (a1x - a2x);
        ru += gv * (a0y - a2y); rv += gv * (a1y - a2y);
        ((float2*)ib.grast)[p] = make_float2(ru, rv);
    }
}

// Buckets of the texel table of a 16 x ROWS region: one per pixel up to 256 (35 KB of LDS), 512
// for larger regions.
constexpr int tex_table_buckets(int rows) { return rows <= 16 ? 256 : 512; }

// Texel-splat region of 16 x ROWS pixels per workgroup (ROWS / 16 pixels per thread): the
// texel grads of the region are summed in one LDS hash table keyed by (level, texel) and flushed
// with one set of atomics per touched texel.  A first pass over the covered pixels' go finds the
// region's fixed-point scale.
template <int ROWS>
struct KTexBwd4TiledF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int W, int H, int tilesX, TexArgs t, TexGradArgs g, const float* rast, const float* uv,
                                                       const float* uvDA, const float* gout, float* guv, float* guvDA, int mode,
                                                       InterpBwdArgs ib) const {
        constexpr int NPX = ROWS / 16;
        __shared__ TexGradTable<tex_table_buckets(ROWS)> tab;
        const bool want_tex = g.tex[0] != nullptr;
        const int view = vb.b.y;
        const int tx = vb.b.x % tilesX, ty = vb.b.x / tilesX;
        // wave w takes the pixels of parity class (w & 1, w >> 1) of each 16 x 16 band, so its lanes
        // sit two pixels apart and fewer of them add to the same texel in one LDS instruction
        const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
        const int lx = 2 * (ln & 7) + (wv & 1), ly = 2 * (ln >> 3) + (wv >> 1);
        const int px = tx * 16 + lx;
        auto pixel = [&](int k) { return ((size_t)view * H + ty * ROWS + k * 16 + ly) * W + px; };
        auto covered = [&](int k) { return px < W && ty * ROWS + k * 16 + ly < H && rast[4 * pixel(k) + 3] >= 1.f; };
        int S = 0;
        bool splat = false, direct = false;
        if (want_tex) {
            tex_table_clear(tab);
            u32 mbits = 0;
#pragma unroll
            for (int k = 0; k < NPX; k++) {
                if (!covered(k)) continue;
                const float4 go = ((const float4*)gout)[pixel(k)];
                mbits = max(mbits, max(max(__float_as_uint(fabsf(go.x)), __float_as_uint(fabsf(go.y))),
                                       max(__float_as_uint(fabsf(go.z)), __float_as_uint(fabsf(go.w)))));
            }
            mbits = wave_umax(mbits);
            if ((threadIdx.x & 63) == 0) tab.wave_max[threadIdx.x >> 6] = mbits;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < BS / 64; w++) mbits = max(mbits, tab.wave_max[w]);
            // all-zero go adds nothing; a non-finite one takes the float path
            splat = mbits != 0;
            direct = mbits >= FIXED_NONFINITE;
            S = fixed_scale(mbits);
        }
#pragma unroll 1
        for (int k = 0; k < NPX; k++) {
            if (!covered(k)) continue;
            const size_t p = pixel(k);
            const float4 go = ((const float4*)gout)[p];
            const float4 gs = make_float4(ldexpf(go.x, S), ldexpf(go.y, S), ldexpf(go.z, S), ldexpf(go.w, S));
            tex_bwd4_pixel(tab, splat, direct, p, go, gs, t, g, uv, uvDA, guv, guvDA, mode, ib);
        }
        if (!splat || direct) return;
        __syncthreads();
        tex_table_flush(tab, g, S);
    }
};
template <int ROWS>
__global__ void __launch_bounds__(BS) k_tex_bwd4_tiled(int W, int H, int tilesX, TexArgs t, TexGradArgs g, const float* rast, const float* uv,
                                                       const float* uvDA, const float* gout, float* guv, float* guvDA, int mode,
                                                       InterpBwdArgs ib) { KTexBwd4TiledF<ROWS>{}(vb_hw(), W, H, tilesX, t, g, rast, uv, uvDA, gout, guv, guvDA, mode, ib); }

// The tiled texture backward (its nt x depth grid flattened) with the SG lobe sums' segment stage
// as workgroups after its own: the two are independent, and the stage's workgroups run in the
// texture backward's tail instead of as a launch of their own.
template <int ROWS>
struct KTexBwdStageF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int nt, int depth, int W, int H, int tilesX, TexArgs t, TexGradArgs g, const float* rast,
                                               const float* uv, const float* uvDA, const float* gout, float* guv, float* guvDA, int mode,
                                               InterpBwdArgs ib, VclColsumStage cs, int nseg) const {
        const u32 ntex = (u32)nt * (u32)depth;
        if (vb.b.x < ntex) {
            KTexBwd4TiledF<ROWS>{}(VB{{vb.b.x % (u32)nt, vb.b.x / (u32)nt, 0u}, {(u32)nt, (u32)depth, 1u}}, W, H, tilesX, t, g, rast, uv, uvDA, gout, guv,
                                   guvDA, mode, ib);
            return;
        }
        const u32 c = vb.b.x - ntex, cx = ((u32)cs.cols + 63u) / 64u;
        KColsumStageF{}(VB{{c % cx, c / cx, 0u}, {cx, (u32)nseg, 1u}}, cs.part, cs.rows, cs.cols, cs.seg_rows, cs.seg);
    }
};
template <int ROWS>
__global__ void __launch_bounds__(BS) k_tex_bwd_stage(int nt, int depth, int W, int H, int tilesX, TexArgs t, TexGradArgs g, const float* rast,
                                                      const float* uv, const float* uvDA, const float* gout, float* guv, float* guvDA, int mode,
                                                      InterpBwdArgs ib, VclColsumStage cs, int nseg) {
    KTexBwdStageF<ROWS>{}(vb_hw(), nt, depth, W, H, tilesX, t, g, rast, uv, uvDA, gout, guv, guvDA, mode, ib, cs, nseg);
}

// ---------------------------------------------------------------- mip utilities
struct KDownsampleF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float* src, float* dst, int w, int h, int C, float* zero_grad) const {
#pragma clang fp contract(off)
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t n = (size_t)w * h * C;
        if (i >= n) return;
        int c = (int)(i % C);
        size_t q = i / C;
        int x = (int)(q % w), y = (int)(q / w);
        int sw = 2 * w;
        float p00 = src[((size_t)(2 * y) * sw + 2 * x) * C + c], p01 = src[((size_t)(2 * y) * sw + 2 * x + 1) * C + c];
        float p10 = src[((size_t)(2 * y + 1) * sw + 2 * x) * C + c], p11 = src[((size_t)(2 * y + 1) * sw + 2 * x + 1) * C + c];
        dst[i] = (p00 + p01 + p10 + p11) * 0.25f;
        if (zero_grad) zero_grad[i] = 0.f;
    }
};
__global__ void __launch_bounds__(BS) k_downsample(const float* src, float* dst, int w, int h, int C, float* zero_grad) { KDownsampleF{}(vb_hw(), src, dst, w, h, C, zero_grad); }

struct KFoldF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float* src, float* dst, int w, int h, int C, float scale) const {
#pragma clang fp contract(off)
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t n = (size_t)w * h * C;
        if (i >= n) return;
        int c = (int)(i % C);
        size_t q = i / C;
        int x = (int)(q % w), y = (int)(q / w);
        int dw = 2 * w;
        float g = src[i] * scale;
        add_nan_check(dst[((size_t)(2 * y) * dw + 2 * x) * C + c], g);
        add_nan_check(dst[((size_t)(2 * y) * dw + 2 * x + 1) * C + c], g);
        add_nan_check(dst[((size_t)(2 * y + 1) * dw + 2 * x) * C + c], g);
        add_nan_check(dst[((size_t)(2 * y + 1) * dw + 2 * x + 1) * C + c], g);
    }
};
__global__ void __launch_bounds__(BS) k_fold(const float* src, float* dst, int w, int h, int C, float scale) { KFoldF{}(vb_hw(), src, dst, w, h, C, scale); }

struct KClampF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, float* x, size_t n, float lo, float hi) const {
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (i >= n) return;
        float v = x[i];
        x[i] = v < lo ? lo : v > hi ? hi : v;
    }
};
__global__ void __launch_bounds__(BS) k_clamp(float* x, size_t n, float lo, float hi) { KClampF{}(vb_hw(), x, n, lo, hi); }

// ---------------------------------------------------------------- Adam
constexpr int ADAM_MAX = 8;
// The loss from the nb per-block partials, by one workgroup of NT threads in a fixed order:
// thread k sums a contiguous run with four interleaved accumulators, then the waves, then the
// wave totals in order.
template <int NT>
__device__ __forceinline__ void mse_final_block(const float* part, size_t nb, float* loss) {
    const size_t per = (nb + NT - 1) / NT;
    const size_t b = threadIdx.x * per, e = min(nb, b + per);
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, r3 = 0.f;
    size_t i = b;
    for (; i + 4 <= e; i += 4) { r0 += part[i]; r1 += part[i + 1]; r2 += part[i + 2]; r3 += part[i + 3]; }
    for (; i < e; i++) r0 += part[i];
    float r = (r0 + r1) + (r2 + r3);
    r = wave_sum(r);
    __shared__ float sh[NT / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < NT / 64; w++) t += sh[w];
        *loss = t * 0.5f;
    }
}

struct AdamArgs {
    int n;
    float* param[ADAM_MAX];
    const float* grad[ADAM_MAX];
    float* m[ADAM_MAX];
    float* v[ADAM_MAX];
    long long size[ADAM_MAX];
    long long block_start[ADAM_MAX + 1];
    int per_thread[ADAM_MAX];   // elements per thread: ADAM_PER_BLOCK / BS, or 1 for tensors whose
                                // elements each sum a column of lobe partials
    double rhom[ADAM_MAX], rhov[ADAM_MAX], rhomt[ADAM_MAX], rhovt[ADAM_MAX], eta[ADAM_MAX], eps[ADAM_MAX];
    const double* bt;   // device {rhomt, rhovt} shared by every tensor (graph replays), or null
    int zero_grad;      // clear each grad after use (the caller's grad clear, fused)
    // tensor 0's grad (positions, 3 floats per vertex) also takes the project backward of pgclip
    // (pviews x pV clip-space grads) through pmats, when pgclip is not null
    const float* pgclip;
    const float* pmats;
    int pviews, pV;
    // the loss's final sum from loss_nb partials, by the launch's last workgroup, when loss is
    // not null
    const float* loss_part;
    size_t loss_nb;
    float* loss;
    // tensors 1..3 (lobe axis, sharpness, amplitude) also take the SG backward's lobe sums:
    // column i * SG_COLS + k of lseg, summed over lnseg segment rows (k_sg_lobe_final's work)
    const float* lseg;
    int lnseg, lK, lC;
};

constexpr int ADAM_SG_COLS = 3 + 1 + 4;  // the SG backward's per-lobe columns (axis, sharpness, C <= 4 amplitudes)

// k_sg_lobe_final's sum for element e of lobe tensor t (1 axis, 2 sharpness, 3 amplitude).
__device__ __forceinline__ float adam_lobe_sum(const AdamArgs& a, int t, long long e) {
    int i, k;
    if (t == 1) { i = (int)(e / 3); k = (int)(e % 3); }
    else if (t == 2) { i = (int)e; k = 3; }
    else { i = (int)(e / a.lC); k = 4 + (int)(e % a.lC); }
    const size_t c = (size_t)i * ADAM_SG_COLS + k, stride = (size_t)a.lK * ADAM_SG_COLS;
    float s = 0.f;
    // unrolled so that the loads run ahead of the (in-order) adds
#pragma unroll 8
    for (int r = 0; r < a.lnseg; r++) s += a.lseg[(size_t)r * stride + c];
    return s;
}

// ProjectBackward's vec grad (project_bwd above) of element i = 3 v + j of the positions.
__device__ __forceinline__ float adam_proj_grad(const AdamArgs& a, long long i) {
    const int v = (int)(i / 3), j = (int)(i % 3);
    float acc = 0.f;
    for (int s = 0; s < a.pviews; s++) {
        const float* m = a.pmats + 16 * s + 4 * j;
        const float4 g = ((const float4*)a.pgclip)[(size_t)s * a.pV + v];
        acc += m[0] * g.x; acc += m[1] * g.y; acc += m[2] * g.z; acc += m[3] * g.w;
    }
    return acc;
}

__device__ __forceinline__ void adam_elem(float& param, float g, float& m, float& v, double rhom, double rhov, double rhomt, double rhovt,
                                          double eta, double eps) {
#pragma clang fp contract(off)
    m = (float)(rhom * (double)m + (1 - rhom) * (double)g);
    v = (float)(rhov * (double)v + (1 - rhov) * (double)g * (double)g);
    double mh = (double)m / (1 - rhomt);
    double vh = (double)v / (1 - rhovt);
    add_nan_check(param, (float)(-mh * eta / (sqrt(vh) + eps)));
}

constexpr int ADAM_PER_BLOCK = BS * 4;

struct KAdamF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, AdamArgs a) const {
        if (a.loss && vb.b.x == vb.g.x - 1) {
            mse_final_block<BS>(a.loss_part, a.loss_nb, a.loss);
            return;
        }
        long long b = vb.b.x;
        int t = 0;
        while (t + 1 < a.n && b >= a.block_start[t + 1]) t++;
        const int ept = a.per_thread[t];
        long long base = (b - a.block_start[t]) * BS * ept;
        float* P = a.param[t]; const float* G = a.grad[t]; float* M = a.m[t]; float* Vv = a.v[t];
        long long n = a.size[t];
        const double rmt = a.bt ? a.bt[0] : a.rhomt[t], rvt = a.bt ? a.bt[1] : a.rhovt[t];
        for (int k = 0; k < ept; k++) {
            long long i = base + k * BS + threadIdx.x;
            if (i >= n) break;
            float p = P[i], m = M[i], v = Vv[i];
            float g = G[i];
            if (t == 0 && a.pgclip) g += adam_proj_grad(a, i);
            if (t >= 1 && t <= 3 && a.lseg) add_nan_check(g, adam_lobe_sum(a, t, i));
            adam_elem(p, g, m, v, a.rhom[t], a.rhov[t], rmt, rvt, a.eta[t], a.eps[t]);
            P[i] = p; M[i] = m; Vv[i] = v;
            if (a.zero_grad) const_cast<float*>(G)[i] = 0.f;
        }
    }
};
__global__ void __launch_bounds__(BS) k_adam(AdamArgs a) { KAdamF{}(vb_hw(), a); }

// Fused texture update on level 0: folded grad (coarse to fine, AddNaNcheck chain of
// gardAddThrough, buffer.cu:615-636) -> Adam -> clamp; level-0 grad zeroed.
struct TexUpdArgs {
    int tw, th, C, L;
    float* tex0;
    float* tex1;   // level 1 (k_tex_update4m writes it)
    float* grad[VCL_TEX_MAX_MIP_LEVEL];
    float* m;
    float* v;
    double rhom, rhov, rhomt, rhovt, eta, eps;
    float lo, hi, scale;
    const double* bt;   // device {rhomt, rhovt} (graph replays), or null
};

// Adam::step bumps rhomt/rhovt before the launch (optimize.cu:109-115); kept on the device so
// a captured step advances them on every replay.
struct KTexUpdateF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, TexUpdArgs a) const {
#pragma clang fp contract(off)
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t n = (size_t)a.tw * a.th * a.C;
        if (i >= n) return;
        int c = (int)(i % a.C);
        size_t q = i / a.C;
        int x = (int)(q % a.tw), y = (int)(q / a.tw);
        float acc = 0.f;
        if (a.L > 1) {
            int l = a.L - 1;
            acc = a.grad[l][((size_t)(y >> l) * (a.tw >> l) + (x >> l)) * a.C + c];
            for (l = a.L - 2; l >= 1; l--) {
                float tmp = a.grad[l][((size_t)(y >> l) * (a.tw >> l) + (x >> l)) * a.C + c];
                add_nan_check(tmp, acc * a.scale);
                acc = tmp;
            }
        }
        float g = a.grad[0][i];
        if (a.L > 1) add_nan_check(g, acc * a.scale);
        float p = a.tex0[i], m = a.m[i], v = a.v[i];
        adam_elem(p, g, m, v, a.rhom, a.rhov, a.bt ? a.bt[0] : a.rhomt, a.bt ? a.bt[1] : a.rhovt, a.eta, a.eps);
        p = p < a.lo ? a.lo : p > a.hi ? a.hi : p;
        a.tex0[i] = p; a.m[i] = m; a.v[i] = v;
        a.grad[0][i] = 0.f;
    }
};
__global__ void __launch_bounds__(BS) k_tex_update(TexUpdArgs a) { KTexUpdateF{}(vb_hw(), a); }

// C == 4: one thread per level-0 texel, float4 traffic; every ancestor grad is loaded up front
// (compile-time-unrolled levels) so the fold costs one memory latency, not one per level.
__device__ __forceinline__ void add_nan_check4(float4& a, float4 b) {
    add_nan_check(a.x, b.x); add_nan_check(a.y, b.y); add_nan_check(a.z, b.z); add_nan_check(a.w, b.w);
}

struct KTexUpdate4F {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, TexUpdArgs a) const {
#pragma clang fp contract(off)
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t n = (size_t)a.tw * a.th;
        if (i >= n) return;
        int x = (int)(i % a.tw), y = (int)(i / a.tw);
        float4 anc[VCL_TEX_MAX_MIP_LEVEL];
#pragma unroll
        for (int l = 1; l < VCL_TEX_MAX_MIP_LEVEL; l++)
            if (l < a.L) anc[l] = ((const float4*)a.grad[l])[(size_t)(y >> l) * (a.tw >> l) + (x >> l)];
        float4 g = ((const float4*)a.grad[0])[i];
        float4 p = ((const float4*)a.tex0)[i], m = ((const float4*)a.m)[i], v = ((const float4*)a.v)[i];
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int l = VCL_TEX_MAX_MIP_LEVEL - 1; l >= 1; l--) {
            if (l == a.L - 1) acc = anc[l];
            else if (l < a.L - 1) {
                float4 t = anc[l];
                add_nan_check4(t, make_float4(acc.x * a.scale, acc.y * a.scale, acc.z * a.scale, acc.w * a.scale));
                acc = t;
            }
        }
        if (a.L > 1) add_nan_check4(g, make_float4(acc.x * a.scale, acc.y * a.scale, acc.z * a.scale, acc.w * a.scale));
        const double rmt = a.bt ? a.bt[0] : a.rhomt, rvt = a.bt ? a.bt[1] : a.rhovt;
        adam_elem(p.x, g.x, m.x, v.x, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
        adam_elem(p.y, g.y, m.y, v.y, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
        adam_elem(p.z, g.z, m.z, v.z, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
        adam_elem(p.w, g.w, m.w, v.w, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
        p.x = p.x < a.lo ? a.lo : p.x > a.hi ? a.hi : p.x;
        p.y = p.y < a.lo ? a.lo : p.y > a.hi ? a.hi : p.y;
        p.z = p.z < a.lo ? a.lo : p.z > a.hi ? a.hi : p.z;
        p.w = p.w < a.lo ? a.lo : p.w > a.hi ? a.hi : p.w;
        ((float4*)a.tex0)[i] = p;
        ((float4*)a.m)[i] = m;
        ((float4*)a.v)[i] = v;
        ((float4*)a.grad[0])[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
};
__global__ void __launch_bounds__(BS) k_tex_update4(TexUpdArgs a) { KTexUpdate4F{}(vb_hw(), a); }

// Lane S's v in every lane of its quad (lanes 4k .. 4k + 3; DPP quad_perm, no LDS).
template <int S>
__device__ __forceinline__ float quad_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), S * 0x55, 0xf, 0xf, false));
}

// k_tex_update4 that also builds level 1 (L > 1, even level-0 sizes): one texel per lane as
// there, with each wave on a 32 x 2 texel patch (a workgroup on 128 x 2) and the four texels of
// a level-1 quad in the lanes of one DPP quad (lane & 1: x offset, lane & 2: y offset).  The
// coarse grads that fold into a texel are the same for the whole level-1 quad, so the quad's
// lane c folds channel c only (one float per level instead of a float4 in every lane) and the
// four results are broadcast in the quad.  After the updates the quad's first lane gathers the
// other three new values and writes their mean in k_downsample4's order, and clears the level-1
// grad (which the quad read for the fold before).
struct KTexUpdate4mF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, TexUpdArgs a) const {
#pragma clang fp contract(off)
        const int lane = threadIdx.x & 63;
        const int c = lane & 3;
        const int x = vb.b.x * 128 + (threadIdx.x >> 6) * 32 + 2 * (lane >> 2) + (lane & 1);
        const int y = vb.b.y * 2 + ((lane >> 1) & 1);
        const bool in = x < a.tw;   // a.tw and a.th are even: a quad is all in or all out
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 p = z4, g = z4, m = z4, v = z4;
        float acc = 0.f;
        const size_t i = (size_t)y * a.tw + x;
        if (in) {
            float anc[VCL_TEX_MAX_MIP_LEVEL];
#pragma unroll
            for (int l = 1; l < VCL_TEX_MAX_MIP_LEVEL; l++)
                if (l < a.L) anc[l] = a.grad[l][((size_t)(y >> l) * (a.tw >> l) + (x >> l)) * 4 + c];
            g = ((const float4*)a.grad[0])[i];
            m = ((const float4*)a.m)[i];
            v = ((const float4*)a.v)[i];
            p = ((const float4*)a.tex0)[i];
#pragma unroll
            for (int l = VCL_TEX_MAX_MIP_LEVEL - 1; l >= 1; l--) {
                if (l == a.L - 1) acc = anc[l];
                else if (l < a.L - 1) {
                    float t = anc[l];
                    add_nan_check(t, acc * a.scale);
                    acc = t;
                }
            }
        }
        const float4 acc4 = make_float4(quad_lane<0>(acc), quad_lane<1>(acc), quad_lane<2>(acc), quad_lane<3>(acc));
        if (in) {
            add_nan_check4(g, make_float4(acc4.x * a.scale, acc4.y * a.scale, acc4.z * a.scale, acc4.w * a.scale));
            const double rmt = a.bt ? a.bt[0] : a.rhomt, rvt = a.bt ? a.bt[1] : a.rhovt;
            adam_elem(p.x, g.x, m.x, v.x, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
            adam_elem(p.y, g.y, m.y, v.y, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
            adam_elem(p.z, g.z, m.z, v.z, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
            adam_elem(p.w, g.w, m.w, v.w, a.rhom, a.rhov, rmt, rvt, a.eta, a.eps);
            p.x = p.x < a.lo ? a.lo : p.x > a.hi ? a.hi : p.x;
            p.y = p.y < a.lo ? a.lo : p.y > a.hi ? a.hi : p.y;
            p.z = p.z < a.lo ? a.lo : p.z > a.hi ? a.hi : p.z;
            p.w = p.w < a.lo ? a.lo : p.w > a.hi ? a.hi : p.w;
            ((float4*)a.tex0)[i] = p;
            ((float4*)a.m)[i] = m;
            ((float4*)a.v)[i] = v;
            ((float4*)a.grad[0])[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const float4 p01 = make_float4(quad_lane<1>(p.x), quad_lane<1>(p.y), quad_lane<1>(p.z), quad_lane<1>(p.w));
        const float4 p10 = make_float4(quad_lane<2>(p.x), quad_lane<2>(p.y), quad_lane<2>(p.z), quad_lane<2>(p.w));
        const float4 p11 = make_float4(quad_lane<3>(p.x), quad_lane<3>(p.y), quad_lane<3>(p.z), quad_lane<3>(p.w));
        if (in && c == 0) {
            const size_t j = (size_t)(y >> 1) * (a.tw >> 1) + (x >> 1);
            ((float4*)a.tex1)[j] = make_float4((((p.x + p01.x) + p10.x) + p11.x) * 0.25f, (((p.y + p01.y) + p10.y) + p11.y) * 0.25f,
                                               (((p.z + p01.z) + p10.z) + p11.z) * 0.25f, (((p.w + p01.w) + p10.w) + p11.w) * 0.25f);
            ((float4*)a.grad[1])[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
};
__global__ void __launch_bounds__(BS) k_tex_update4m(TexUpdArgs a) { KTexUpdate4mF{}(vb_hw(), a); }

// C == 4 mip level from the level above (float4 per texel), clearing that level's grads.
struct KDownsample4F {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float4* src, float4* dst, int w, int h, float4* zero_grad) const {
#pragma clang fp contract(off)
        size_t i = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (i >= (size_t)w * h) return;
        int x = (int)(i % w), y = (int)(i / w);
        size_t sw = 2 * (size_t)w;
        float4 p00 = src[(2 * y) * sw + 2 * x], p01 = src[(2 * y) * sw + 2 * x + 1];
        float4 p10 = src[(2 * y + 1) * sw + 2 * x], p11 = src[(2 * y + 1) * sw + 2 * x + 1];
        dst[i] = make_float4((((p00.x + p01.x) + p10.x) + p11.x) * 0.25f, (((p00.y + p01.y) + p10.y) + p11.y) * 0.25f,
                             (((p00.z + p01.z) + p10.z) + p11.z) * 0.25f, (((p00.w + p01.w) + p10.w) + p11.w) * 0.25f);
        if (zero_grad) zero_grad[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
};
__global__ void __launch_bounds__(BS) k_downsample4(const float4* src, float4* dst, int w, int h, float4* zero_grad) { KDownsample4F{}(vb_hw(), src, dst, w, h, zero_grad); }

// Several mip levels in one launch: each workgroup reads a BSZ x BSZ block of level s (BSZ = 64
// or 32) and writes the block's texels of levels s+1 .. s+nl (BSZ/2 squared down to 1^2, nl at
// most log2(BSZ)), each from the previous level kept in LDS with the same box-filter
// arithmetic as k_downsample4 (so the values are identical), clearing the matching coarse
// grads.  32 x 32 blocks give a large first level four times the workgroups.
struct MipArgs {
    const float4* src;
    int sw, sh, nl;
    float4* dst[6];
    float4* zg[6];
};

template <int BSZ>
struct KMipFusedF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, MipArgs a) const {
#pragma clang fp contract(off)
        __shared__ float4 buf[2][(BSZ / 2) * (BSZ / 2)];
        const int bx = vb.b.x, by = vb.b.y;
        for (int k = 1; k <= a.nl; k++) {
            const int bs = BSZ >> k;                     // block side at level s+k
            const int w = a.sw >> k, h = a.sh >> k;      // level size
            const int ox = bx * bs, oy = by * bs;
            float4* out = buf[k & 1];
            const float4* in = buf[(k - 1) & 1];
            for (int i = threadIdx.x; i < bs * bs; i += BS) {
                const int lx = i % bs, ly = i / bs, x = ox + lx, y = oy + ly;
                if (x >= w || y >= h) continue;
                float4 p00, p01, p10, p11;
                if (k == 1) {
                    const size_t sw = (size_t)a.sw;
                    p00 = a.src[(2 * y) * sw + 2 * x]; p01 = a.src[(2 * y) * sw + 2 * x + 1];
                    p10 = a.src[(2 * y + 1) * sw + 2 * x]; p11 = a.src[(2 * y + 1) * sw + 2 * x + 1];
                } else {
                    const int pb = 2 * bs;
                    p00 = in[(2 * ly) * pb + 2 * lx]; p01 = in[(2 * ly) * pb + 2 * lx + 1];
                    p10 = in[(2 * ly + 1) * pb + 2 * lx]; p11 = in[(2 * ly + 1) * pb + 2 * lx + 1];
                }
                const float4 v = make_float4((((p00.x + p01.x) + p10.x) + p11.x) * 0.25f, (((p00.y + p01.y) + p10.y) + p11.y) * 0.25f,
                                             (((p00.z + p01.z) + p10.z) + p11.z) * 0.25f, (((p00.w + p01.w) + p10.w) + p11.w) * 0.25f);
                out[ly * bs + lx] = v;
                a.dst[k - 1][(size_t)y * w + x] = v;
                if (a.zg[k - 1]) a.zg[k - 1][(size_t)y * w + x] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            __syncthreads();
        }
    }
};
template <int BSZ>
__global__ void __launch_bounds__(BS) k_mip_fused(MipArgs a) { KMipFusedF<BSZ>{}(vb_hw(), a); }

// The Adam launch with one more workgroup that runs the mip rebuild's last pass when that pass
// is a single workgroup (the coarsest levels, from a source level of at most 64 x 64 texels),
// which then costs no launch of its own.
struct KAdamMipF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, AdamArgs a, MipArgs m) const {
        if (vb.b.x == vb.g.x - 1) {
            KMipFusedF<64>{}(VB{{0u, 0u, 0u}, {1u, 1u, 1u}}, m);
            return;
        }
        KAdamF{}(VB{vb.b, {vb.g.x - 1, vb.g.y, vb.g.z}}, a);
    }
};
__global__ void __launch_bounds__(BS) k_adam_mip(AdamArgs a, MipArgs m) { KAdamMipF{}(vb_hw(), a, m); }

// ---------------------------------------------------------------- NormalAxis / ViewAxis
// NormalAxis of one pixel: interpolated vertex normal, normalised, rotated (zero off-mesh).
// The covered pixel's axis from its triangle's three vertex normals n0, n1, n2.
__device__ __forceinline__ void normal_axis_eval(float4 r, const float* rot, float3 n0, float3 n1, float3 n2, float& o0, float& o1, float& o2) {
#pragma clang fp contract(off)
    float N0 = (n0.x - n2.x) * r.x + (n1.x - n2.x) * r.y + n2.x;
    float N1 = (n0.y - n2.y) * r.x + (n1.y - n2.y) * r.y + n2.y;
    float N2 = (n0.z - n2.z) * r.x + (n1.z - n2.z) * r.y + n2.z;
    float d = 1.f / sqrtf(N0 * N0 + N1 * N1 + N2 * N2);
    N0 *= d; N1 *= d; N2 *= d;
    o0 = rot[0] * N0 + rot[4] * N1 + rot[8] * N2;
    o1 = rot[1] * N0 + rot[5] * N1 + rot[9] * N2;
    o2 = rot[2] * N0 + rot[6] * N1 + rot[10] * N2;
}

__device__ __forceinline__ float3 load_f3(const float* a, u32 i) {
    const float* q = a + 3 * (size_t)i;
    return make_float3(q[0], q[1], q[2]);
}

__device__ __forceinline__ void normal_axis_px(float4 r, const float* rot, const float* normal, const u32* nidx, float& o0, float& o1, float& o2) {
    int id = (int)r.w - 1;
    o0 = 0.f; o1 = 0.f; o2 = 0.f;
    if (id >= 0)
        normal_axis_eval(r, rot, load_f3(normal, nidx[3 * id]), load_f3(normal, nidx[3 * id + 1]), load_f3(normal, nidx[3 * id + 2]), o0, o1, o2);
}

struct KNormalAxisF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, const float* rast, const float* rot, const float* normal, const u32* nidx, float* out) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float o0, o1, o2;
        normal_axis_px(((const float4*)rast)[p], rot, normal, nidx, o0, o1, o2);
        out[3 * p] = o0; out[3 * p + 1] = o1; out[3 * p + 2] = o2;
    }
};
__global__ void __launch_bounds__(BS) k_normal_axis(size_t P, const float* rast, const float* rot, const float* normal, const u32* nidx, float* out) { KNormalAxisF{}(vb_hw(), P, rast, rot, normal, nidx, out); }

struct Mat3x8 { float m[8][9]; };

// ViewAxis of one pixel from its layer's 3x3 inverse projection-view block m.
__device__ __forceinline__ void view_axis_px(int W, int H, int px, int py, const float* m, float& o0, float& o1, float& o2) {
    float sx = (float)px / (float)W * 2.f - 1.f, sy = (float)py / (float)H * 2.f - 1.f;
    float v0 = m[0] * sx + m[3] * sy + m[6];
    float v1 = m[1] * sx + m[4] * sy + m[7];
    float v2 = m[2] * sx + m[5] * sy + m[8];
    float d = 1.f / sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
    o0 = v0 * d; o1 = v1 * d; o2 = v2 * d;
}

struct KViewAxisF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int W, int H, int depth, const float* pvinv_dev, Mat3x8 pv_small, float* out) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t P = (size_t)W * H * depth;
        if (p >= P) return;
        int px = (int)(p % W), py = (int)((p / W) % H), s = (int)(p / ((size_t)W * H));
        const float* m = pvinv_dev ? pvinv_dev + 9 * s : pv_small.m[s];
        view_axis_px(W, H, px, py, m, out[3 * p], out[3 * p + 1], out[3 * p + 2]);
    }
};
__global__ void __launch_bounds__(BS) k_view_axis(int W, int H, int depth, const float* pvinv_dev, Mat3x8 pv_small, float* out) { KViewAxisF{}(vb_hw(), W, H, depth, pvinv_dev, pv_small, out); }

// ---------------------------------------------------------------- SGSpecular
// (float)(x / fmax((double)y, 1e-6)) for float x, y, as the reference writes it: where y is at or
// above the smallest float not below 1e-6 the quotient of the two floats in double, rounded to
// float, is the correctly rounded float quotient (double carries more than twice float's 24 bits),
// so a float division gives the same bits; below it (or for a NaN y) the double form runs.
__device__ __forceinline__ float div_fmax_1e6(float x, float y) {
    return y >= 0x1.0c6f7cp-20f ? x / y : (float)((double)x / 1e-6);
}

__device__ __forceinline__ void spec_terms(float nx, float ny, float nz, float vx_, float vy_, float vz_, float m, float ior, float* p, float& amp,
                                           float& sharp, float* dA, float* dS) {
    float vx = -vx_, vy = -vy_, vz = -vz_;
    float vdn = vx * nx + vy * ny + vz * nz;
    p[0] = 2.f * vdn * nx - vx; p[1] = 2.f * vdn * ny - vy; p[2] = 2.f * vdn * nz - vz;
    float m2 = m * m;
    float vn = fmaxf(vdn, 0.f);
    float vn2 = vn * vn;
    float iGv = sqrtf(vn2 + m2 * (1.f - vn2));
    float Gv = div_fmax_1e6(1.f, vn + iGv);
    float pdn = p[0] * nx + p[1] * ny + p[2] * nz;
    float pn = fmaxf(pdn, 0.f);
    float pn2 = pn * pn;
    float iGp = sqrtf(pn2 + m2 * (1.f - pn2));
    float Gp = div_fmax_1e6(1.f, pn + iGp);
    float g = sqrtf(ior * ior - 1.f + pn2);
    float gp = g + pn, gm = g - pn;
    float f0 = gm / gp;
    float f1 = (pn * gp - 1.f) / (pn * gm + 1.f);
    float F = f0 * f0 * (1.f + f1 * f1);
    amp = div_fmax_1e6(F * Gp * Gv * pn, m2);
    sharp = div_fmax_1e6(0.5f, m2 * vn);
    if (dA) {
        float dGvdm = (float)(-m * (1.f - vn2) / fmax((double)iGv, 1e-6) * Gv * Gv);
        float dGpdm = (float)(-m * (1.f - pn2) / fmax((double)iGp, 1e-6) * Gp * Gp);
        *dA = (float)(F * pn * (-2.0 * Gp * Gv / fmax((double)(m2 * m), 1e-6) + (Gp * dGvdm + Gv * dGpdm) / fmax((double)m2, 1e-6)));
        *dS = (float)(-1.0 / fmax((double)(m2 * m * vn), 1e-6));
    }
}

struct KSgSpecFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, const float* rast, const float* normal, const float* view, const float* rough,
                                                    float ior, float* oAxis, float* oSharp, float* oAmp) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float pa[3] = {0.f, 0.f, 0.f}, amp = 0.f, sh = 0.f;
        if ((int)rast[4 * p + 3] - 1 >= 0)
            spec_terms(normal[3 * p], normal[3 * p + 1], normal[3 * p + 2], view[3 * p], view[3 * p + 1], view[3 * p + 2], rough[p], ior, pa, amp, sh,
                       nullptr, nullptr);
        oAxis[3 * p] = pa[0]; oAxis[3 * p + 1] = pa[1]; oAxis[3 * p + 2] = pa[2];
        oSharp[p] = sh; oAmp[p] = amp;
    }
};
__global__ void __launch_bounds__(BS) k_sg_spec_fwd(size_t P, const float* rast, const float* normal, const float* view, const float* rough,
                                                    float ior, float* oAxis, float* oSharp, float* oAmp) { KSgSpecFwdF{}(vb_hw(), P, rast, normal, view, rough, ior, oAxis, oSharp, oAmp); }

// The fit step's NormalAxis -> ViewAxis -> SGSpecular chain in one pass per pixel: the same
// per-pixel code as the three ops, without the round trip of the normal and view axes through
// memory between them.
struct KShadePrepF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int W, int H, int depth, const float* rast, const float* rot, const float* normal, const u32* nidx,
                                                   const float* pvinv, const float* rough, float ior, float* onrm, float* oview, float* oAxis,
                                                   float* oSharp, float* oAmp) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        size_t P = (size_t)W * H * depth;
        if (p >= P) return;
        const int px = (int)(p % W), py = (int)((p / W) % H), s = (int)(p / ((size_t)W * H));
        const float4 r = ((const float4*)rast)[p];
        float n0, n1, n2, v0, v1, v2;
        normal_axis_px(r, rot, normal, nidx, n0, n1, n2);
        view_axis_px(W, H, px, py, pvinv + 9 * s, v0, v1, v2);
        onrm[3 * p] = n0; onrm[3 * p + 1] = n1; onrm[3 * p + 2] = n2;
        oview[3 * p] = v0; oview[3 * p + 1] = v1; oview[3 * p + 2] = v2;
        float pa[3] = {0.f, 0.f, 0.f}, amp = 0.f, sh = 0.f;
        if ((int)r.w - 1 >= 0) spec_terms(n0, n1, n2, v0, v1, v2, rough[p], ior, pa, amp, sh, nullptr, nullptr);
        oAxis[3 * p] = pa[0]; oAxis[3 * p + 1] = pa[1]; oAxis[3 * p + 2] = pa[2];
        oSharp[p] = sh; oAmp[p] = amp;
    }
};
__global__ void __launch_bounds__(BS) k_shade_prep(int W, int H, int depth, const float* rast, const float* rot, const float* normal, const u32* nidx,
                                                   const float* pvinv, const float* rough, float ior, float* onrm, float* oview, float* oAxis,
                                                   float* oSharp, float* oAmp) { KShadePrepF{}(vb_hw(), W, H, depth, rast, rot, normal, nidx, pvinv, rough, ior, onrm, oview, oAxis, oSharp, oAmp); }

// The rasterizer's per-pixel resolve as the fit step runs it inside its per-pixel pass: the
// winning-key buffer, the clip-space vertices (V per view) and the outputs rast / db.
struct ResolveArgs {
    const unsigned long long* keys;
    const float* proj;
    int V;
    float xs, ys;
    float* rast;
    float* db;
};

// The fit step's per-pixel passes ahead of the SG forward in one (C == 4, normals indexed like
// the uv attributes): the rasterizer's resolve (a pixel's winning key -> rast and db, as
// raster.hip's k_resolve), uv interpolation with derivatives and the texture sample (the
// arithmetic of KInterpTex4F) and the shading inputs NormalAxis, ViewAxis and SGSpecular (that of
// KShadePrepF).  They share the triangle's index loads, the clip-space vertices and the uv and
// normal attributes are loaded together, and the specular terms are formed while the texel
// loads are in flight.  Each thread takes NPX pixels one workgroup-width apart, each stage's
// loads for all its pixels issued together.
template <int NPX>
struct KPixelPrep4F {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, int W, int H, int depth, ResolveArgs ra, const float* uvattr, const u32* idx,
                                               float* uvp, float* uvDA, TexArgs t, float* out, const float* rot, const float* normal,
                                               const float* pvinv, const float* rough, float ior, float* onrm, float* oview, float* oAxis,
                                               float* oSharp, float* oAmp) const {
        const size_t WH = (size_t)W * H, P = WH * depth;
        const size_t p0 = (size_t)vb.b.x * (BS * NPX) + threadIdx.x;
        unsigned long long key[NPX];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            key[j] = p < P ? ra.keys[p] : RAST_NO_KEY;
        }
        u32 i0[NPX], i1[NPX], i2[NPX];
        float rg[NPX];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            i0[j] = i1[j] = i2[j] = 0;
            rg[j] = 0.f;
            if (key[j] != RAST_NO_KEY) {
                const u32 tri = (u32)(key[j] & 0xffffffffu);
                i0[j] = idx[3 * tri]; i1[j] = idx[3 * tri + 1]; i2[j] = idx[3 * tri + 2];
                rg[j] = rough[p0 + (size_t)j * BS];
            }
        }
        float4 r[NPX], b[NPX];
        float2 a0[NPX], a1[NPX], a2[NPX];
        float3 n0[NPX], n1[NPX], n2[NPX];
        {
            float4 c0[NPX], c1[NPX], c2[NPX];
#pragma unroll
            for (int j = 0; j < NPX; j++) {
                a0[j] = a1[j] = a2[j] = make_float2(0.f, 0.f);
                n0[j] = n1[j] = n2[j] = make_float3(0.f, 0.f, 0.f);
                c0[j] = c1[j] = c2[j] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (key[j] != RAST_NO_KEY) {
                    const float4* clip = (const float4*)ra.proj + (size_t)((p0 + (size_t)j * BS) / WH) * ra.V;
                    c0[j] = clip[i0[j]]; c1[j] = clip[i1[j]]; c2[j] = clip[i2[j]];
                    a0[j] = ((const float2*)uvattr)[i0[j]];
                    a1[j] = ((const float2*)uvattr)[i1[j]];
                    a2[j] = ((const float2*)uvattr)[i2[j]];
                    n0[j] = load_f3(normal, i0[j]);
                    n1[j] = load_f3(normal, i1[j]);
                    n2[j] = load_f3(normal, i2[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < NPX; j++) {
                const size_t p = p0 + (size_t)j * BS;
                r[j] = b[j] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (key[j] != RAST_NO_KEY) {
                    const int px = (int)(p % W), py = (int)((p / W) % H);
                    resolve_pixel(c0[j], c1[j], c2[j], (u32)(key[j] & 0xffffffffu), ra.xs, ra.ys, pixel_centre(px, ra.xs), pixel_centre(py, ra.ys), true,
                                  r[j], b[j]);
                }
                if (p < P) {
                    ((float4*)ra.rast)[p] = r[j];
                    ((float4*)ra.db)[p] = b[j];
                }
            }
        }
        Tex4Taps q[NPX];
        float4 tx[NPX][8];
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            float2 u = make_float2(0.f, 0.f);
            float4 da = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((int)r[j].w - 1 >= 0) {
                const double w2 = 1.0 - (double)r[j].x - (double)r[j].y;
                float2 dau, dav;
                interp_attr_px(r[j], b[j], w2, a0[j].x, a1[j].x, a2[j].x, u.x, dau);
                interp_attr_px(r[j], b[j], w2, a0[j].y, a1[j].y, a2[j].y, u.y, dav);
                da = make_float4(dau.x, dau.y, dav.x, dav.y);
            }
            if (p < P) {
                ((float2*)uvp)[p] = u;
                ((float4*)uvDA)[p] = da;
            }
            if (r[j].w >= 1.f) {
                q[j] = tex4_taps(t, u, da);
                tx[j][0] = q[j].T0[q[j].i0.x]; tx[j][1] = q[j].T0[q[j].i0.y]; tx[j][2] = q[j].T0[q[j].i0.z]; tx[j][3] = q[j].T0[q[j].i0.w];
                if (q[j].two) {
                    tx[j][4] = q[j].T1[q[j].i1.x]; tx[j][5] = q[j].T1[q[j].i1.y]; tx[j][6] = q[j].T1[q[j].i1.z]; tx[j][7] = q[j].T1[q[j].i1.w];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            if (p >= P) continue;
            const int px = (int)(p % W), py = (int)((p / W) % H), s = (int)(p / WH);
            float nx = 0.f, ny = 0.f, nz = 0.f, vx, vy, vz;
            if ((int)r[j].w - 1 >= 0) normal_axis_eval(r[j], rot, n0[j], n1[j], n2[j], nx, ny, nz);
            view_axis_px(W, H, px, py, pvinv + 9 * s, vx, vy, vz);
            onrm[3 * p] = nx; onrm[3 * p + 1] = ny; onrm[3 * p + 2] = nz;
            if (oview) { oview[3 * p] = vx; oview[3 * p + 1] = vy; oview[3 * p + 2] = vz; }
            float pa[3] = {0.f, 0.f, 0.f}, amp = 0.f, sh = 0.f;
            if ((int)r[j].w - 1 >= 0) spec_terms(nx, ny, nz, vx, vy, vz, rg[j], ior, pa, amp, sh, nullptr, nullptr);
            oAxis[3 * p] = pa[0]; oAxis[3 * p + 1] = pa[1]; oAxis[3 * p + 2] = pa[2];
            oSharp[p] = sh; oAmp[p] = amp;
        }
#pragma unroll
        for (int j = 0; j < NPX; j++) {
            const size_t p = p0 + (size_t)j * BS;
            if (p >= P) continue;
            float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r[j].w >= 1.f) {
                o = bilerp4(tx[j][0], tx[j][1], tx[j][2], tx[j][3], q[j].fx, q[j].fy);
                if (q[j].two) o = lerp4(o, bilerp4(tx[j][4], tx[j][5], tx[j][6], tx[j][7], q[j].gx, q[j].gy), q[j].f);
            }
            ((float4*)out)[p] = o;
        }
    }
};
constexpr int PIXEL_PREP_NPX = 2;
__global__ void __launch_bounds__(BS) k_pixel_prep4(int W, int H, int depth, ResolveArgs ra, const float* uvattr, const u32* idx, float* uvp,
                                                    float* uvDA, TexArgs t, float* out, const float* rot, const float* normal, const float* pvinv,
                                                    const float* rough, float ior, float* onrm, float* oview, float* oAxis, float* oSharp,
                                                    float* oAmp) {
    KPixelPrep4F<PIXEL_PREP_NPX>{}(vb_hw(), W, H, depth, ra, uvattr, idx, uvp, uvDA, t, out, rot, normal, pvinv, rough, ior, onrm, oview, oAxis,
                                   oSharp, oAmp);
}

struct KSgSpecBwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, const float* rast, const float* normal, const float* view, const float* rough,
                                                    float ior, const float* gAmp, const float* gSharp, float* grough) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        if ((int)rast[4 * p + 3] - 1 < 0) return;
        float pa[3], amp, sh, dA, dS;
        spec_terms(normal[3 * p], normal[3 * p + 1], normal[3 * p + 2], view[3 * p], view[3 * p + 1], view[3 * p + 2], rough[p], ior, pa, amp, sh, &dA,
                   &dS);
        float g = grough[p];
        add_nan_check(g, gAmp[p] * dA);
        add_nan_check(g, gSharp[p] * dS);
        grough[p] = g;
    }
};
__global__ void __launch_bounds__(BS) k_sg_spec_bwd(size_t P, const float* rast, const float* normal, const float* view, const float* rough,
                                                    float ior, const float* gAmp, const float* gSharp, float* grough) { KSgSpecBwdF{}(vb_hw(), P, rast, normal, view, rough, ior, gAmp, gSharp, grough); }

// ---------------------------------------------------------------- SphericalGaussian
constexpr float SG_DS = 2.133f, SG_DA = 2.34f;
constexpr int SG_MAXC = 4;

__device__ __forceinline__ float sg_g(float l, float c, float amp) {
    return (float)((expf(l - c) - expf(-l - c)) / fmax((double)l, 1e-6) * amp);
}

struct KSgFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int C, int K, const float* rast, const float* __restrict__ axis, const float* __restrict__ sharp,
                                               const float* __restrict__ amp, const float* normal, const float* diffuse, const float* sAxis,
                                               const float* sSharp, const float* sAmp, float* out, float* denv, float* senv) const {
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float de[SG_MAXC] = {0.f, 0.f, 0.f, 0.f}, se[SG_MAXC] = {0.f, 0.f, 0.f, 0.f};
        bool cov = (int)rast[4 * p + 3] - 1 >= 0;
        if (cov) {
            float dnx = normal[3 * p] * SG_DS, dny = normal[3 * p + 1] * SG_DS, dnz = normal[3 * p + 2] * SG_DS;
            float ss = sSharp[p], sa = sAmp[p];
            float snx = sAxis[3 * p] * ss, sny = sAxis[3 * p + 1] * ss, snz = sAxis[3 * p + 2] * ss;
            for (int i = 0; i < K; i++) {
                float lam = sharp[i];
                float gx = axis[3 * i] * lam, gy = axis[3 * i + 1] * lam, gz = axis[3 * i + 2] * lam;
                float wx = dnx + gx, wy = dny + gy, wz = dnz + gz;
                float dl = sqrtf(wx * wx + wy * wy + wz * wz);
                float dg = sg_g(dl, SG_DS + lam, SG_DA);
                float qx = snx + gx, qy = sny + gy, qz = snz + gz;
                float sl = sqrtf(qx * qx + qy * qy + qz * qz);
                float sg = sg_g(sl, ss + lam, sa);
                for (int k = 0; k < C; k++) {
                    float a = amp[i * C + k];
                    add_nan_check(de[k], dg * a);
                    add_nan_check(se[k], sg * a);
                }
            }
        }
        for (int k = 0; k < C; k++) {
            out[p * C + k] = cov ? diffuse[p * C + k] * de[k] + se[k] : 0.f;
            denv[p * C + k] = de[k];
            senv[p * C + k] = se[k];
        }
    }
};
__global__ void __launch_bounds__(BS) k_sg_fwd(size_t P, int C, int K, const float* rast, const float* __restrict__ axis, const float* __restrict__ sharp,
                                               const float* __restrict__ amp, const float* normal, const float* diffuse, const float* sAxis,
                                               const float* sSharp, const float* sAmp, float* out, float* denv, float* senv) { KSgFwdF{}(vb_hw(), P, C, K, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out, denv, senv); }

// Backward: each block owns SG_PIX_PER_THREAD * 256 pixels per sweep and a fixed set of
// sweeps; lobe grads are summed over its pixels (wave reductions, then LDS) and written as
// one row of partials; rows are summed in a fixed order afterwards.
constexpr int SG_NPIX = 2;
constexpr int SG_COLS = 3 + 1 + SG_MAXC;  // axis, sharpness, amplitude
static_assert(SG_COLS == ADAM_SG_COLS, "the Adam launch's lobe-sum columns are the SG backward's");

__global__ void __launch_bounds__(BS) k_sg_bwd(size_t P, int C, int K, const float* rast, const float* __restrict__ axis,
                                               const float* __restrict__ sharp, const float* __restrict__ amp, const float* normal,
                                               const float* diffuse, const float* sAxis, const float* sSharp, const float* sAmp, const float* gout,
                                               float* gdiffuse, float* gnormal, float* gsAxis, float* gsSharp, float* gsAmp, float* part,
                                               int want_lobes) {
    extern __shared__ float lds[];  // [4 waves][K][SG_COLS], only when want_lobes
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (want_lobes) {
        for (int i = threadIdx.x; i < 4 * K * SG_COLS; i += BS) lds[i] = 0.f;
        __syncthreads();
    }
    size_t sweep = (size_t)BS * SG_NPIX;
    for (size_t base = (size_t)blockIdx.x * sweep; base < P; base += (size_t)gridDim.x * sweep) {
        // per-pixel state in registers
        bool cov[SG_NPIX];
        size_t pp[SG_NPIX];
        float dn[SG_NPIX][3], sn[SG_NPIX][3], ss_[SG_NPIX], sa_[SG_NPIX], go[SG_NPIX][SG_MAXC], df[SG_NPIX][SG_MAXC];
        float gn[SG_NPIX][3], gp[SG_NPIX][3], gss[SG_NPIX], gsa[SG_NPIX], de[SG_NPIX][SG_MAXC];
#pragma unroll
        for (int j = 0; j < SG_NPIX; j++) {
            size_t p = base + (size_t)j * BS + threadIdx.x;
            pp[j] = p;
            cov[j] = p < P && (int)rast[4 * p + 3] - 1 >= 0;
            for (int k = 0; k < 3; k++) { gn[j][k] = 0.f; gp[j][k] = 0.f; dn[j][k] = 0.f; sn[j][k] = 0.f; }
            gss[j] = 0.f; gsa[j] = 0.f; ss_[j] = 0.f; sa_[j] = 0.f;
            for (int k = 0; k < SG_MAXC; k++) { go[j][k] = 0.f; df[j][k] = 0.f; de[j][k] = 0.f; }
            if (cov[j]) {
                for (int k = 0; k < 3; k++) dn[j][k] = normal[3 * p + k] * SG_DS;
                ss_[j] = sSharp[p]; sa_[j] = sAmp[p];
                for (int k = 0; k < 3; k++) sn[j][k] = sAxis[3 * p + k] * ss_[j];
                for (int k = 0; k < C; k++) { go[j][k] = gout[p * C + k]; df[j][k] = diffuse[p * C + k]; }
            }
        }
        for (int i = 0; i < K; i++) {
            float lam = sharp[i];
            float mx = axis[3 * i], my = axis[3 * i + 1], mz = axis[3 * i + 2];
            float gx = mx * lam, gy = my * lam, gz = mz * lam;
            float aa[SG_MAXC];
            for (int k = 0; k < C; k++) aa[k] = amp[i * C + k];
            float acc[SG_COLS];
            for (int k = 0; k < SG_COLS; k++) acc[k] = 0.f;
#pragma unroll
            for (int j = 0; j < SG_NPIX; j++) {
                if (!cov[j]) continue;
                float wx = dn[j][0] + gx, wy = dn[j][1] + gy, wz = dn[j][2] + gz;
                float l = sqrtf(wx * wx + wy * wy + wz * wz);
                float il = (float)(1.0 / fmax((double)l, 1e-6));
                float e1 = expf(l - SG_DS - lam), e2 = expf(-l - SG_DS - lam);
                float dgv = (e1 - e2) * il;
                float dgdl = ((1.f - il) * e1 + (1.f + il) * e2) * il;
                float qx = sn[j][0] + gx, qy = sn[j][1] + gy, qz = sn[j][2] + gz;
                float sl = sqrtf(qx * qx + qy * qy + qz * qz);
                float isl = (float)(1.0 / fmax((double)sl, 1e-6));
                float f1 = expf(sl - ss_[j] - lam), f2 = expf(-sl - ss_[j] - lam);
                float sgv = (f1 - f2) * isl;
                float sgdl = ((1.f - isl) * f1 + (1.f + isl) * f2) * isl;
                float sdiff = 0.f, ssum = 0.f;
                float dgf = (float)((e1 - e2) / fmax((double)l, 1e-6) * SG_DA);
                for (int k = 0; k < C; k++) {
                    float g = go[j][k];
                    sdiff += g * aa[k] * df[j][k];
                    ssum += g * aa[k];
                    acc[4 + k] += g * (dgv * SG_DA * df[j][k] + sgv * sa_[j]);
                    add_nan_check(de[j][k], dgf * aa[k]);
                }
                float cd = sdiff * SG_DA * dgdl * il, cs = ssum * sa_[j] * sgdl * isl;
                float vx = cd * wx + cs * qx, vy = cd * wy + cs * qy, vz = cd * wz + cs * qz;
                acc[0] += vx * lam; acc[1] += vy * lam; acc[2] += vz * lam;
                float mdw = mx * wx + my * wy + mz * wz, mdq = mx * qx + my * qy + mz * qz;
                acc[3] += sdiff * SG_DA * (dgdl * mdw * il - dgv) + ssum * sa_[j] * (sgdl * mdq * isl - sgv);
                gn[j][0] += cd * wx * SG_DS; gn[j][1] += cd * wy * SG_DS; gn[j][2] += cd * wz * SG_DS;
                gp[j][0] += cs * qx * ss_[j]; gp[j][1] += cs * qy * ss_[j]; gp[j][2] += cs * qz * ss_[j];
                float pdq = (sn[j][0] * qx + sn[j][1] * qy + sn[j][2] * qz) / (ss_[j] != 0.f ? ss_[j] : 1.f);
                gss[j] += ssum * sa_[j] * (sgdl * pdq * isl - sgv);
                gsa[j] += ssum * sgv;
            }
            if (want_lobes) {
                for (int k = 0; k < 4 + C; k++) {
                    float v = wave_sum(acc[k]);
                    if (lane == 0) lds[(wid * K + i) * SG_COLS + k] += v;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < SG_NPIX; j++) {
            if (!cov[j]) continue;
            size_t p = pp[j];
            if (gdiffuse) for (int k = 0; k < C; k++) add_nan_check(gdiffuse[p * C + k], go[j][k] * de[j][k]);
            if (gnormal) for (int k = 0; k < 3; k++) gnormal[3 * p + k] += gn[j][k];
            if (gsAxis) for (int k = 0; k < 3; k++) gsAxis[3 * p + k] += gp[j][k];
            if (gsSharp) gsSharp[p] += gss[j];
            if (gsAmp) gsAmp[p] += gsa[j];
        }
    }
    if (!want_lobes) return;
    __syncthreads();
    for (int i = threadIdx.x; i < K * SG_COLS; i += BS) {
        float s = ((lds[i] + lds[K * SG_COLS + i]) + lds[2 * K * SG_COLS + i]) + lds[3 * K * SG_COLS + i];
        part[(size_t)blockIdx.x * K * SG_COLS + i] = s;
    }
}

// Final stage of the lobe-grad column sums: column i * SG_COLS + k (lobe i) is the sum of the
// segment partials in segment order, added (AddNaNcheck) into the axis (k < 3), sharpness
// (k == 3) or amplitude (k >= 4) grad of lobe i.
struct KSgLobeFinalF {
    static constexpr int kBS = 256, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float* seg, int nseg, int K, int C, float* gaxis, float* gsharp, float* gamp) const {
        const int c = vb.b.x * blockDim.x + threadIdx.x;
        if (c >= K * SG_COLS) return;
        float s = 0.f;
#pragma unroll 8
        for (int r = 0; r < nseg; r++) s += seg[(size_t)r * K * SG_COLS + c];
        const int i = c / SG_COLS, k = c % SG_COLS;
        if (k < 3) { if (gaxis) add_nan_check(gaxis[3 * i + k], s); }
        else if (k == 3) { if (gsharp) add_nan_check(gsharp[i], s); }
        else if (k - 4 < C && gamp) add_nan_check(gamp[i * C + (k - 4)], s);
    }
};
__global__ void __launch_bounds__(256) k_sg_lobe_final(const float* seg, int nseg, int K, int C, float* gaxis, float* gsharp, float* gamp) { KSgLobeFinalF{}(vb_hw(), seg, nseg, K, C, gaxis, gsharp, gamp); }

// ---- fast SG kernels: native exp2 / sqrt / rcp (1-ulp hardware ops), lobes staged in LDS.
constexpr float L2E = 1.4426950408889634f;
constexpr int SG_MAXK = 256;

__device__ __forceinline__ float nexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float nsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float nrcp(float x) { return __builtin_amdgcn_rcpf(x); }
// AddNaNcheck for an accumulator that is finite by construction.
__device__ __forceinline__ void acc_checked(float& a, float b) { float v = a + b; a = isfinite(v) ? v : a; }

struct SgLobe { float gx, gy, gz, lam, mx, my, mz, lamL2E, a[4], L2; };  // lamL2E = lam * log2(e), L2 = |lam mu|^2

__device__ __forceinline__ void sg_stage_lobes(SgLobe* L, int K, int C, const float* axis, const float* sharp, const float* amp) {
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        float lam = sharp[i];
        SgLobe l;
        l.mx = axis[3 * i]; l.my = axis[3 * i + 1]; l.mz = axis[3 * i + 2];
        l.gx = l.mx * lam; l.gy = l.my * lam; l.gz = l.mz * lam; l.lam = lam; l.lamL2E = lam * L2E;
        l.L2 = l.gx * l.gx + l.gy * l.gy + l.gz * l.gz;
        for (int k = 0; k < 4; k++) l.a[k] = k < C ? amp[i * C + k] : 0.f;
        L[i] = l;
    }
}

// One lobe term: g = (e^{l-c} - e^{-l-c}) / max(l, 1e-6), plus (optionally) dg/dl.
__device__ __forceinline__ float sg_term(float wx, float wy, float wz, float base, float lam, float& l, float& il, float& e1, float& e2) {
    l = nsqrt(wx * wx + wy * wy + wz * wz);
    il = nrcp(fmaxf(l, 1e-6f));
    e1 = nexp2(((l - base) - lam) * L2E);
    e2 = nexp2(((-l - base) - lam) * L2E);
    return (e1 - e2) * il;
}

struct KSgFwdFastF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int C, int K, const float* rast, const float* axis, const float* sharp,
                                                    const float* amp, const float* normal, const float* diffuse, const float* sAxis,
                                                    const float* sSharp, const float* sAmp, float* out, float* denv, float* senv) const {
        __shared__ SgLobe lobes[SG_MAXK];
        sg_stage_lobes(lobes, K, C, axis, sharp, amp);
        __syncthreads();
        size_t p = (size_t)vb.b.x * blockDim.x + threadIdx.x;
        if (p >= P) return;
        float de[4] = {0.f, 0.f, 0.f, 0.f}, se[4] = {0.f, 0.f, 0.f, 0.f};
        bool cov = (int)rast[4 * p + 3] - 1 >= 0;
        if (cov) {
            float dnx = normal[3 * p] * SG_DS, dny = normal[3 * p + 1] * SG_DS, dnz = normal[3 * p + 2] * SG_DS;
            float ss = sSharp[p], sa = sAmp[p];
            float snx = sAxis[3 * p] * ss, sny = sAxis[3 * p + 1] * ss, snz = sAxis[3 * p + 2] * ss;
            for (int i = 0; i < K; i++) {
                const SgLobe& L = lobes[i];
                float l, il, e1, e2;
                float dg = sg_term(dnx + L.gx, dny + L.gy, dnz + L.gz, SG_DS, L.lam, l, il, e1, e2) * SG_DA;
                float sg = sg_term(snx + L.gx, sny + L.gy, snz + L.gz, ss, L.lam, l, il, e1, e2) * sa;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (k < C) { acc_checked(de[k], dg * L.a[k]); acc_checked(se[k], sg * L.a[k]); }
                }
            }
        }
        if (C == 4) {
            ((float4*)out)[p] = cov ? make_float4(diffuse[4 * p] * de[0] + se[0], diffuse[4 * p + 1] * de[1] + se[1], diffuse[4 * p + 2] * de[2] + se[2],
                                                  diffuse[4 * p + 3] * de[3] + se[3])
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
            ((float4*)denv)[p] = make_float4(de[0], de[1], de[2], de[3]);
            ((float4*)senv)[p] = make_float4(se[0], se[1], se[2], se[3]);
        } else {
            for (int k = 0; k < C; k++) {
                out[p * C + k] = cov ? diffuse[p * C + k] * de[k] + se[k] : 0.f;
                denv[p * C + k] = de[k];
                senv[p * C + k] = se[k];
            }
        }
    }
};
__global__ void __launch_bounds__(BS) k_sg_fwd_fast(size_t P, int C, int K, const float* rast, const float* axis, const float* sharp,
                                                    const float* amp, const float* normal, const float* diffuse, const float* sAxis,
                                                    const float* sSharp, const float* sAmp, float* out, float* denv, float* senv) { KSgFwdFastF{}(vb_hw(), P, C, K, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out, denv, senv); }

// C == 4 forward.  A workgroup takes 1024 (or 512) consecutive pixels and compacts the covered
// ones, so its waves shade covered pixels only (a thread-per-pixel wave at a silhouette runs the whole
// lobe loop for a few covered lanes).  Per lobe: two 16-byte LDS reads ((lam mu, cD) and the
// amplitude) and one 4-byte read (lam log2 e).
// AddNaNcheck per channel reduces to one finiteness test per term: the term is bounded
// (e^{l-c} <= 1 since l <= c, and 1/max(l, 1e-6) <= 1e6), so a finite term times finite
// amplitudes always passes the per-channel check and a non-finite one always fails it; a block
// whose lobes hold a non-finite amplitude takes the per-channel path.  1/l and l come from one
// rsq.  e^{-l-c} is dropped when it is below half an ulp of e^{l-c} for every lane of the wave
// (l > 8.75 gives e^{-2l} < 2^-25), from l = |w + lam*mu| >= | |lam*mu| - |w| | with |w| the
// pixel's 2.133*normal or spec-sharpness*spec-axis length (both lengths carry a relative
// margin for their float rounding).
// Bounded form: when every lobe of the workgroup and every pixel of the wave is finite with
// |lam mu| - lam <= 8, |w| - base <= 8, amplitudes within 2^32 and sharpness within 2^20 (the
// normalised axes and normals of a fit), each term's exponent l - c is at most 16 (17 with its
// float rounding) and 1/l at most 1e6, so a term times its amplitudes is below
// e^17 * 1e6 * 2^32 * 2^32 < 2^109 and every partial sum of at most 256 of them is finite: no
// term can fail AddNaNcheck, the loop tests none, and the
// constant factors (2.34, the spec amplitude) multiply the two lobe sums once at the end
// instead of every term.
// g = (lam mu, cD), cD = (SG_DS + lam) log2(e), lamL2E = lam log2(e), len2 = |lam mu|^2, len = |lam mu|
struct SgLobe4 { float4 g; float4 a; float lamL2E, len2, len, lam; };

// (e^{l-c} - e^{-l-c}) / l with the exponent's constant part c = (base + lam) log2(e) given: one
// rsq gives 1/l and l, each exponent is one fma (within 1e-6 relative of the reference's
// (l - c) - lam form), and e^{-l-c} is formed only when `both` (wave-uniform).
__device__ __forceinline__ float sg_term_rsq(float wx, float wy, float wz, float c, bool both) {
    float d = wx * wx + wy * wy + wz * wz;
    float il = __builtin_amdgcn_rsqf(fmaxf(d, 1e-12f));
    float l = d * il;
    float t = nexp2(fmaf(l, L2E, -c));
    if (both) {
        // the empty asm keeps this a branch the wave skips rather than an exp computed on
        // every lobe and selected away
        asm volatile("" :::);
        t = t - nexp2(fmaf(-l, L2E, -c));
    }
    return t * il;
}

// sg_term_rsq for the bounded form, from the squared length d = |w|^2 formed as
// (|p|^2 + |lam mu|^2) + 2 p . lam mu (one add and three fma instead of the three adds, one
// multiply and two fma of |p + lam mu|^2); within the bounds both squares are finite.
__device__ __forceinline__ float sg_term_d(float d, float c, bool both) {
    float il = __builtin_amdgcn_rsqf(fmaxf(d, 1e-12f));
    float l = d * il;
    float t = nexp2(fmaf(l, L2E, -c));
    if (both) {
        asm volatile("" :::);
        t = t - nexp2(fmaf(-l, L2E, -c));
    }
    return t * il;
}

// e^{-l-c} is below half an ulp of e^{l-c} once l > SG_E2_DROP (e^{-17.5} < 2^-25).
constexpr float SG_E2_DROP = 8.75f;
// Bounded form's limits: |lam mu| - lam and |w| - base at most SG_FWD_EXCESS, amplitudes and
// spec amplitudes below SG_FWD_AMAX in magnitude, lobe and spec sharpness at most SG_FWD_LMAX
// in magnitude (so the squared lengths are finite and the exponents' float rounding, about
// ulp(lam log2(e)), stays below 1).
constexpr float SG_FWD_EXCESS = 8.f, SG_FWD_AMAX = 4294967296.f, SG_FWD_LMAX = 1048576.f;

// Which lobes of chunk [c0, c0 + 64) may need e^{-l-c} for some lane of the wave: lane j tests
// lobe c0 + j against the wave's range [lmin, lmax] of the pixel-side length |w| (the per-lane
// test |len - |w|| <= SG_E2_DROP + 0.01 + 1e-5 (len + |w|) implies len within that range widened
// by the largest threshold), and the ballot gives the chunk's mask (a superset of the lobes the
// per-lane test selects; an exponential formed where it was not needed is below half an ulp and
// changes nothing).  All 64 lanes must call it.
__device__ __forceinline__ u64 sg_e2_mask(const SgLobe4* lobes, int K, int c0, float lmin, float lmax) {
    const int j = c0 + (int)(threadIdx.x & 63);
    const float gl = j < K ? lobes[j].len : __builtin_nanf("");
    const float T = SG_E2_DROP + 0.01f + 1e-5f * (gl + lmax);
    return __ballot(gl >= lmin - T && gl <= lmax + T);
}

__device__ __forceinline__ float wave_fmin(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_fmax(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Forms of the lobe loop: per-channel AddNaNcheck (a workgroup whose lobes hold a non-finite
// amplitude), one finiteness test per term, or the bounded form (no test; 2.34 and the spec
// amplitude applied to the sums afterwards).  The choice is uniform per wave, so the forms are
// separate loops rather than a branch inside one.
enum SgFwdForm { SG_FWD_CHECKED, SG_FWD_TESTED, SG_FWD_BOUNDED };

// The lobe loop of one pixel.  dmask / smask: per 64-lobe chunk, the lobes whose diffuse /
// specular e^{-l-c} is formed (sg_e2_mask), tested as scalars.
template <SgFwdForm FORM>
__device__ __forceinline__ void sg_fwd4_lobes(const SgLobe4* lobes, int K, float dnx, float dny, float dnz, float snx, float sny, float snz,
                                              float ss, float sa, const u64* dmask, const u64* smask, float4& de, float4& se) {
    const float ssL2E = ss * L2E;
    // bounded form: 2 p and |p|^2 of the two pixel-side vectors
    const float dnx2 = 2.f * dnx, dny2 = 2.f * dny, dnz2 = 2.f * dnz, dsq = dnx * dnx + dny * dny + dnz * dnz;
    const float snx2 = 2.f * snx, sny2 = 2.f * sny, snz2 = 2.f * snz, ssq = snx * snx + sny * sny + snz * snz;
#pragma unroll
    for (int c = 0; c < SG_MAXK / 64; c++) {
        if (c * 64 >= K) break;
        const u64 dm = dmask[c], sm = smask[c];
        const int iend = min(K, c * 64 + 64);
#pragma unroll 2
        for (int i = c * 64; i < iend; i++) {
            const float4 g = lobes[i].g, a = lobes[i].a;
            const bool dboth = (dm >> (i - c * 64)) & 1ull;
            const bool sboth = (sm >> (i - c * 64)) & 1ull;
            if (FORM == SG_FWD_BOUNDED) {
                // (lam log2 e, |lam mu|^2) as one 8-byte read: two 4-byte reads, scheduled apart,
                // each had the loop form the lobe's LDS address again (1024^2: 95 -> 91 us)
                const float2 ll = *reinterpret_cast<const float2*>(&lobes[i].lamL2E);
                const float L = ll.y;
                const float dd = fmaf(dnx2, g.x, fmaf(dny2, g.y, fmaf(dnz2, g.z, dsq + L)));
                const float ds = fmaf(snx2, g.x, fmaf(sny2, g.y, fmaf(snz2, g.z, ssq + L)));
                const float td = sg_term_d(dd, g.w, dboth);
                const float ts = sg_term_d(ds, ssL2E + ll.x, sboth);
                de.x += td * a.x; de.y += td * a.y; de.z += td * a.z; de.w += td * a.w;
                se.x += ts * a.x; se.y += ts * a.y; se.z += ts * a.z; se.w += ts * a.w;
                continue;
            }
            const float td = sg_term_rsq(dnx + g.x, dny + g.y, dnz + g.z, g.w, dboth);
            const float ts = sg_term_rsq(snx + g.x, sny + g.y, snz + g.z, ssL2E + lobes[i].lamL2E, sboth);
            if (FORM == SG_FWD_TESTED) {
                const float dg = td * SG_DA, sg = ts * sa;
                const float dv = isfinite(dg) ? dg : 0.f, sv = isfinite(sg) ? sg : 0.f;
                de.x += dv * a.x; de.y += dv * a.y; de.z += dv * a.z; de.w += dv * a.w;
                se.x += sv * a.x; se.y += sv * a.y; se.z += sv * a.z; se.w += sv * a.w;
            } else {
                const float dg = td * SG_DA, sg = ts * sa;
                acc_checked(de.x, dg * a.x); acc_checked(de.y, dg * a.y); acc_checked(de.z, dg * a.z); acc_checked(de.w, dg * a.w);
                acc_checked(se.x, sg * a.x); acc_checked(se.y, sg * a.y); acc_checked(se.z, sg * a.z); acc_checked(se.w, sg * a.w);
            }
        }
    }
    if (FORM == SG_FWD_BOUNDED) {
        de.x *= SG_DA; de.y *= SG_DA; de.z *= SG_DA; de.w *= SG_DA;
        se.x *= sa; se.y *= sa; se.z *= sa; se.w *= sa;
    }
}

// Pixels per k_sg_fwd4 workgroup (SG_FWD_PIX) and threads (NT): 1,024 pixels on 256 threads, or
// 512 pixels on 512 threads when the 1,024-pixel grid would hold fewer than 4,096 workgroups.
// With few workgroups all are resident at once and the launch ends with few waves per SIMD
// left to hide the lobe loop's latencies; 512 one-round threads per 512 pixels double the waves
// (two rounds of residency), where 512 pixels on 256 threads left the CUs under-filled
// (1024^2: 104 -> 95 us).  Large grids keep 256 threads (the 512-thread form was 8 % slower at
// 64 x 512^2).
template <int SG_FWD_PIX, int NT>
struct KSgFwd4F {
    static constexpr int kBS = NT, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int K, const float* rast, const float* axis, const float* sharp, const float* amp,
                                                const float* normal, const float* diffuse, const float* sAxis, const float* sSharp,
                                                const float* sAmp, float* out, float* denv, float* senv) const {
        __shared__ SgLobe4 lobes[SG_MAXK];
        __shared__ int nonfinite, unbounded;
        __shared__ u32 plist[SG_FWD_PIX];
        __shared__ u32 wcount[NT / 64];
        if (threadIdx.x == 0) { nonfinite = 0; unbounded = 0; }
        __syncthreads();
        for (int i = threadIdx.x; i < K; i += blockDim.x) {
            float lam = sharp[i];
            float4 a = make_float4(amp[4 * i], amp[4 * i + 1], amp[4 * i + 2], amp[4 * i + 3]);
            float gx = axis[3 * i] * lam, gy = axis[3 * i + 1] * lam, gz = axis[3 * i + 2] * lam;
            const float len = sqrtf(gx * gx + gy * gy + gz * gz);
            lobes[i].g = make_float4(gx, gy, gz, (SG_DS + lam) * L2E);
            lobes[i].a = a;
            lobes[i].len = len;
            lobes[i].len2 = gx * gx + gy * gy + gz * gz;
            lobes[i].lam = lam;
            lobes[i].lamL2E = lam * L2E;
            if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(a.w))) nonfinite = 1;
            // NaN lengths or sharpness fail the comparison and so count as unbounded
            const float amax = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
            if (!(len - lam <= SG_FWD_EXCESS && fabsf(lam) <= SG_FWD_LMAX && amax < SG_FWD_AMAX)) unbounded = 1;
        }
        __syncthreads();
        const bool checked = nonfinite != 0;
        const bool lobes_bounded = unbounded == 0;
        // uncovered pixels get their zeros here; the covered ones of the workgroup's SG_FWD_PIX are
        // compacted (in pixel order) so that every wave shades covered pixels only
        const size_t base = (size_t)vb.b.x * SG_FWD_PIX;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        u32 n = 0;
        for (int k = 0; k < SG_FWD_PIX / NT; k++) {
            const u32 local = (u32)(k * NT + threadIdx.x);
            const size_t p = base + local;
            const bool in = p < P;
            const bool cov = in && (int)rast[4 * p + 3] - 1 >= 0;
            if (in && !cov) { ((float4*)out)[p] = z4; ((float4*)denv)[p] = z4; ((float4*)senv)[p] = z4; }
            block_append(cov, local, plist, wcount, n);
        }
        for (u32 r = 0; r < n; r += NT) {
            const u32 i = r + threadIdx.x;
            const bool act = i < n;
            const size_t p = act ? base + plist[i] : 0;
            float dnx = 0.f, dny = 0.f, dnz = 0.f, ss = 0.f, sa = 0.f, snx = 0.f, sny = 0.f, snz = 0.f;
            float dlen = __builtin_nanf(""), slen = __builtin_nanf("");   // lanes without a pixel: ignored by fmin/fmax
            bool bounded = true;
            if (act) {
                dnx = normal[3 * p] * SG_DS; dny = normal[3 * p + 1] * SG_DS; dnz = normal[3 * p + 2] * SG_DS;
                ss = sSharp[p]; sa = sAmp[p];
                snx = sAxis[3 * p] * ss; sny = sAxis[3 * p + 1] * ss; snz = sAxis[3 * p + 2] * ss;
                dlen = sqrtf(dnx * dnx + dny * dny + dnz * dnz);
                slen = sqrtf(snx * snx + sny * sny + snz * snz);
                bounded = dlen - SG_DS <= SG_FWD_EXCESS && slen - ss <= SG_FWD_EXCESS && fabsf(ss) <= SG_FWD_LMAX && fabsf(sa) < SG_FWD_AMAX;
            }
            const bool wave_bounded = lobes_bounded && __ballot(!bounded) == 0ull;
            const float dmin = wave_fmin(dlen), dmax = wave_fmax(dlen), smin = wave_fmin(slen), smax = wave_fmax(slen);
            u64 dmask[SG_MAXK / 64], smask[SG_MAXK / 64];
#pragma unroll
            for (int c = 0; c < SG_MAXK / 64; c++) {
                dmask[c] = sg_e2_mask(lobes, K, c * 64, dmin, dmax);
                smask[c] = sg_e2_mask(lobes, K, c * 64, smin, smax);
            }
            if (act) {
                float4 de = z4, se = z4;
                if (checked) sg_fwd4_lobes<SG_FWD_CHECKED>(lobes, K, dnx, dny, dnz, snx, sny, snz, ss, sa, dmask, smask, de, se);
                else if (wave_bounded) sg_fwd4_lobes<SG_FWD_BOUNDED>(lobes, K, dnx, dny, dnz, snx, sny, snz, ss, sa, dmask, smask, de, se);
                else sg_fwd4_lobes<SG_FWD_TESTED>(lobes, K, dnx, dny, dnz, snx, sny, snz, ss, sa, dmask, smask, de, se);
                const float4 df = ((const float4*)diffuse)[p];
                ((float4*)out)[p] = make_float4(df.x * de.x + se.x, df.y * de.y + se.y, df.z * de.z + se.z, df.w * de.w + se.w);
                ((float4*)denv)[p] = de;
                ((float4*)senv)[p] = se;
            }
        }
    }
};
template <int SG_FWD_PIX, int NT>
__global__ void __launch_bounds__(NT) k_sg_fwd4(size_t P, int K, const float* rast, const float* axis, const float* sharp, const float* amp,
                                                const float* normal, const float* diffuse, const float* sAxis, const float* sSharp,
                                                const float* sAmp, float* out, float* denv, float* senv) { KSgFwd4F<SG_FWD_PIX, NT>{}(vb_hw(), P, K, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out, denv, senv); }

// Lobe-parallel backward for the lobe gradients (axis, sharpness, amplitude).  Thread t owns
// lobe t % KP for pixel sub-group t / KP; covered pixels of a 256-pixel chunk are compacted into
// LDS (deterministic order), every wave walks them in the same order (LDS broadcast), and the
// eight lobe sums stay in registers across all chunks of the block.  grad.diffuse is
// gout * diffenv (the forward's lobe-ordered sum) written in the staging pass.
// X = gout * diffuse * 2.34 and G = gout * spec amplitude: the constant factors of the diffuse
// and specular terms are taken into the pixel's staged values, so the lobe loop never applies
// them.
// ssL2E = spec sharpness * log2(e); dsq, ssq: the squared lengths of (dnx, dny, dnz) and (snx, sny, snz)
struct __align__(16) SgPix { float dnx, dny, dnz, ssL2E, snx, sny, snz, dsq, X[4], G[4], ssq, pad[3]; };

// One lobe term of the backward: g = (e1 - e2) il and dg/dl = ((e1 + e2) - g) il, with e2 =
// e^{-l-c} formed only when BOTH (otherwise e2 is below half an ulp of e1 and both sums equal e1).
template <bool BOTH>
__device__ __forceinline__ float sg_term_bwd(float wx, float wy, float wz, float c, float& il, float& dgdl) {
    float d = wx * wx + wy * wy + wz * wz;
    il = __builtin_amdgcn_rsqf(fmaxf(d, 1e-12f));
    float l = d * il;
    float e1 = nexp2(fmaf(l, L2E, -c));
    if (BOTH) {
        float e2 = nexp2(fmaf(-l, L2E, -c));
        float g = (e1 - e2) * il;
        dgdl = ((e1 + e2) - g) * il;
        return g;
    }
    float g = e1 * il;
    dgdl = (e1 - g) * il;
    return g;
}

// sg_term_bwd from the squared length d = |w|^2.
template <bool BOTH>
__device__ __forceinline__ float sg_term_bwd_d(float d, float c, float& il, float& dgdl) {
    il = __builtin_amdgcn_rsqf(fmaxf(d, 1e-12f));
    float l = d * il;
    float e1 = nexp2(fmaf(l, L2E, -c));
    if (BOTH) {
        float e2 = nexp2(fmaf(-l, L2E, -c));
        float g = (e1 - e2) * il;
        dgdl = ((e1 + e2) - g) * il;
        return g;
    }
    float g = e1 * il;
    dgdl = (e1 - g) * il;
    return g;
}

// The pixel loop of one lobe lane (see k_sg_bwd_lobes); DB / SB: the diffuse / specular e2 is
// needed by some pixel of the chunk for some lobe of the wave.  The chunk's pixel list is padded
// with zero pixels to a multiple of G, so every lane runs n / G iterations (a zero pixel adds
// zeros: its X and G are 0).  A = 2.34 * sum(gout diffuse amp), B = sa * sum(gout amp).
// The squared lengths |p + lam mu|^2 are formed as (|p|^2 + |lam mu|^2) + 2 p . lam mu, and the
// chain-factor sums over w = p + lam mu as sum(cd p) plus lam mu times sum(cd) (added by the
// caller): the vectors w and q themselves are never formed.
template <bool DB, bool SB>
__device__ __forceinline__ void sg_bwd_lobe_loop(const SgLobe& L, const SgPix* pix, int grp, int nit, int G, float accW[3], float& accC,
                                                 float accX[4], float accG[4]) {
    const float cD = L.lamL2E + SG_DS * L2E;
    const float g2x = 2.f * L.gx, g2y = 2.f * L.gy, g2z = 2.f * L.gz;
    // the record is read as five 16-byte words
    const float4* xp = reinterpret_cast<const float4*>(pix + grp);
    for (int t = 0; t < nit; t++, xp += 5 * G) {
        const float4 d = xp[0], sn = xp[1], X = xp[2], Gp = xp[3];
        const float ssq = xp[4].x;
        const float A = ((X.x * L.a[0] + X.y * L.a[1]) + X.z * L.a[2]) + X.w * L.a[3];
        const float B = ((Gp.x * L.a[0] + Gp.y * L.a[1]) + Gp.z * L.a[2]) + Gp.w * L.a[3];
        float il, dgdl;
        const float dd = fmaf(d.x, g2x, fmaf(d.y, g2y, fmaf(d.z, g2z, sn.w + L.L2)));
        const float dgv = sg_term_bwd_d<DB>(dd, cD, il, dgdl);
        float isl, sgdl;
        const float ds = fmaf(sn.x, g2x, fmaf(sn.y, g2y, fmaf(sn.z, g2z, ssq + L.L2)));
        const float sgv = sg_term_bwd_d<SB>(ds, L.lamL2E + d.w, isl, sgdl);
        accX[0] += X.x * dgv; accX[1] += X.y * dgv; accX[2] += X.z * dgv; accX[3] += X.w * dgv;
        accG[0] += Gp.x * sgv; accG[1] += Gp.y * sgv; accG[2] += Gp.z * sgv; accG[3] += Gp.w * sgv;
        const float cd = A * dgdl * il, cs = B * sgdl * isl;
        accW[0] += cd * d.x; accW[0] += cs * sn.x;
        accW[1] += cd * d.y; accW[1] += cs * sn.y;
        accW[2] += cd * d.z; accW[2] += cs * sn.z;
        accC += cd; accC += cs;
    }
}

// One pixel's inputs to the lobe backward (channels beyond C are 0; gd: the current diffuse grad,
// de: diffenv, loaded when the diffuse grad is wanted; gd_fresh: the diffuse grad holds nothing
// yet, so gd is 0 without a load -- the value a cleared buffer would give).
struct SgRaw { float go[4], df[4], de[4], gd[4], nx, ny, nz, ss, sa, ax, ay, az; bool cov; };

__device__ __forceinline__ void sg_load_raw(SgRaw& r, size_t p, bool cov, int C, const float* gout, const float* diffuse, const float* denv,
                                            const float* gdiffuse, const float* normal, const float* sAxis, const float* sSharp,
                                            const float* sAmp, bool gd_fresh) {
    r.cov = cov;
    if (!cov) return;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        r.go[k] = k < C ? gout[p * C + k] : 0.f;
        r.df[k] = k < C ? diffuse[p * C + k] : 0.f;
        r.de[k] = k < C && gdiffuse ? denv[p * C + k] : 0.f;
        r.gd[k] = k < C && gdiffuse && !gd_fresh ? gdiffuse[p * C + k] : 0.f;
    }
    r.nx = normal[3 * p]; r.ny = normal[3 * p + 1]; r.nz = normal[3 * p + 2];
    r.ss = sSharp[p]; r.sa = sAmp[p];
    r.ax = sAxis[3 * p]; r.ay = sAxis[3 * p + 1]; r.az = sAxis[3 * p + 2];
}

// Does a pixel-side vector of length w need e^{-l-c} for some lobe of length in [lo, hi]?
// l = |w_vec + lobe| >= |len - w|, and e2 drops when l > SG_E2_DROP; the margin covers the
// float rounding of l.  NaN lengths answer yes.
__device__ __forceinline__ bool sg_needs_e2(float w, float lo, float hi) {
    const float T = SG_E2_DROP + 0.01f + 1e-5f * (hi + w);
    return !(w + T < lo) && !(w - T > hi);
}

struct KSgBwdLobesF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t P, int C, int K, int KP, const float* rast, const float* axis, const float* sharp,
                                                     const float* amp, const float* normal, const float* diffuse, const float* sAxis,
                                                     const float* sSharp, const float* sAmp, const float* gout, const float* denv, float* gdiffuse,
                                                     float* part, int want_lobes, int gd_fresh) const {
        __shared__ SgLobe lobes[SG_MAXK];
        // the chunk's pixel records, and after the last chunk the per-thread sums (one buffer)
        static_assert(sizeof(SgPix) * BS >= sizeof(float) * BS * SG_COLS, "reduction fits the pixel buffer");
        __shared__ float4 sbuf[sizeof(SgPix) * BS / 16];
        SgPix* pix = reinterpret_cast<SgPix*>(sbuf);
        float(*red)[SG_COLS] = reinterpret_cast<float(*)[SG_COLS]>(sbuf);
        __shared__ int wcount[BS / 64];
        __shared__ u32 wneed[BS / 64];
        __shared__ unsigned long long skey[SG_MAXK];
        __shared__ int lorig[SG_MAXK];
        __shared__ float2 lrange[BS / 64];
        int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
        // Lane layout: with KP >= 4, wave w holds the KP / 4 lobes [w KP / 4, (w + 1) KP / 4) and
        // 256 / KP pixel groups (lane = group * KP / 4 + lobe offset), so a wave's lobes are a quarter
        // of the staged order; with KP < 4 every wave holds all lobes.  Either way a lobe's pixels
        // go to groups 0 .. G - 1 (G = BS / KP) in the same order.
        const int LPW = KP >= 4 ? KP / 4 : KP;
        int li = KP >= 4 ? wid * LPW + lane % LPW : tid % KP, grp = KP >= 4 ? lane / LPW : tid / KP, G = BS / KP;
        // Lobes are staged in order of |axis * sharpness| (ties by index), so that a wave's 64 lobe
        // lanes hold similar lengths and the e2 test below can pass for a whole wave; the per-lobe
        // sums do not depend on which lane holds a lobe.
        for (int i = tid; i < K; i += BS) {
            const float lam = sharp[i];
            const float gx = axis[3 * i] * lam, gy = axis[3 * i + 1] * lam, gz = axis[3 * i + 2] * lam;
            const float len = sqrtf(gx * gx + gy * gy + gz * gz);
            const u32 kb = isnan(len) ? 0xffffffffu : __float_as_uint(len);
            skey[i] = ((unsigned long long)kb << 32) | (u32)i;
        }
        __syncthreads();
        for (int i = tid; i < K; i += BS) {
            const unsigned long long ki = skey[i];
            int r = 0;
            for (int j = 0; j < K; j++) r += skey[j] < ki;
            const float lam = sharp[i];
            SgLobe l;
            l.mx = axis[3 * i]; l.my = axis[3 * i + 1]; l.mz = axis[3 * i + 2];
            l.gx = l.mx * lam; l.gy = l.my * lam; l.gz = l.mz * lam; l.lam = lam; l.lamL2E = lam * L2E;
            l.L2 = l.gx * l.gx + l.gy * l.gy + l.gz * l.gz;
            for (int k = 0; k < 4; k++) l.a[k] = k < C ? amp[i * C + k] : 0.f;
            lobes[r] = l;
            lorig[r] = i;
        }
        __syncthreads();
        // lobe-length range of each wave's lobes (one range, lw = 0, when KP < 4)
        const int nlw = KP >= 4 ? BS / 64 : 1, lw = KP >= 4 ? wid : 0;
        {
            float len = __builtin_nanf("");
            if (li < K) { const SgLobe& L = lobes[li]; len = sqrtf(L.gx * L.gx + L.gy * L.gy + L.gz * L.gz); }
            const float lo = wave_fmin(len), hi = wave_fmax(len);
            if (lane == 0 && wid < nlw) lrange[lw] = make_float2(lo, hi);
        }
        __syncthreads();
        float acc[SG_COLS];
#pragma unroll
        for (int k = 0; k < SG_COLS; k++) acc[k] = 0.f;
        float accW[3] = {0.f, 0.f, 0.f}, accX[4] = {0.f, 0.f, 0.f, 0.f}, accG[4] = {0.f, 0.f, 0.f, 0.f}, accC = 0.f;
        // The pixel inputs of the next chunk are loaded while this chunk's lobe loop runs (and the
        // coverage word of the chunk after that), so the chunks' loads overlap the lobe work instead
        // of every workgroup waiting on them between chunks.
        // The chunks form bands of nb (the grid size); in band k the workgroup takes position
        // (b + k s) mod nb, s = 2 nb / 3, so its chunks move through the rows' column ranges and
        // through the image's row bands rather than all lying in one column range, where the
        // covered share differs most (for a centred disk covering 66 % of a 1024^2 image the
        // fullest workgroup holds 1.08x the mean covered pixels, against 1.52x in grid-stride
        // order).
        const size_t nbk = vb.g.x, sft = nbk / 3 * 2;
        auto cbase = [&](size_t k) -> size_t { return (k * nbk + (vb.b.x + k * sft) % nbk) * BS; };
        const size_t nband = (P + nbk * BS - 1) / (nbk * BS);
        SgRaw cur, nxt;
        float rwn;
        {
            const size_t p0 = cbase(0) + tid, p1 = cbase(1) + tid;
            sg_load_raw(cur, p0, p0 < P && (int)rast[4 * p0 + 3] - 1 >= 0, C, gout, diffuse, denv, gdiffuse, normal, sAxis, sSharp, sAmp,
                        gd_fresh != 0);
            rwn = p1 < P ? rast[4 * p1 + 3] : 0.f;
        }
        for (size_t kb = 0; kb < nband; kb++) {
            const size_t p = cbase(kb) + tid;
            const bool cov = cur.cov;
            {
                const size_t pn = cbase(kb + 1) + tid, p2 = cbase(kb + 2) + tid;
                sg_load_raw(nxt, pn, pn < P && (int)rwn - 1 >= 0, C, gout, diffuse, denv, gdiffuse, normal, sAxis, sSharp, sAmp,
                            gd_fresh != 0);
                rwn = p2 < P ? rast[4 * p2 + 3] : 0.f;
            }
            SgPix q;
            u32 need = 0;   // bit 2 lw: diffuse e2 for lobe wave lw, bit 2 lw + 1: specular
            if (cov) {
                if (gdiffuse)
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (k < C) {
                            float t = cur.gd[k];
                            add_nan_check(t, cur.go[k] * cur.de[k]);
                            gdiffuse[p * C + k] = t;
                        }
                q.dnx = cur.nx * SG_DS; q.dny = cur.ny * SG_DS; q.dnz = cur.nz * SG_DS;
                const float ss = cur.ss;
                q.ssL2E = ss * L2E;
                q.snx = cur.ax * ss; q.sny = cur.ay * ss; q.snz = cur.az * ss;
                q.dsq = q.dnx * q.dnx + q.dny * q.dny + q.dnz * q.dnz;
                q.ssq = q.snx * q.snx + q.sny * q.sny + q.snz * q.snz;
                q.pad[0] = q.pad[1] = q.pad[2] = 0.f;
#pragma unroll
                for (int k = 0; k < 4; k++) { q.X[k] = cur.go[k] * cur.df[k] * SG_DA; q.G[k] = cur.go[k] * cur.sa; }
                const float dlen = sqrtf(q.dsq);
                const float slen = sqrtf(q.ssq);
                for (int w = 0; w < nlw; w++) {
                    const float2 r = lrange[w];
                    need |= (sg_needs_e2(dlen, r.x, r.y) ? 1u : 0u) << (2 * w);
                    need |= (sg_needs_e2(slen, r.x, r.y) ? 2u : 0u) << (2 * w);
                }
            }
            if (!want_lobes) { cur = nxt; continue; }
            // deterministic compaction: wave ballot + prefix over waves
            unsigned long long m = __ballot(cov);
            int wpos = __popcll(m & ((1ull << lane) - 1ull));
            u32 wn = 0;
            for (int b = 0; b < 2 * nlw; b++) wn |= __ballot((need >> b) & 1u) ? 1u << b : 0u;
            if (lane == 0) { wcount[wid] = __popcll(m); wneed[wid] = wn; }
            __syncthreads();
            int off = 0, ncov = 0;
            u32 cneed = 0;
            for (int w = 0; w < BS / 64; w++) { if (w < wid) off += wcount[w]; ncov += wcount[w]; cneed |= wneed[w]; }
            if (cov) pix[off + wpos] = q;
            // zero pixels up to a multiple of G (G divides BS, so the padded count stays <= BS)
            const int ncovp = (ncov + G - 1) / G * G;
            if (tid >= ncov && tid < ncovp) {
                SgPix z;
                z.dnx = z.dny = z.dnz = z.ssL2E = z.snx = z.sny = z.snz = z.dsq = z.ssq = 0.f;
                z.pad[0] = z.pad[1] = z.pad[2] = 0.f;
#pragma unroll
                for (int k = 0; k < 4; k++) { z.X[k] = 0.f; z.G[k] = 0.f; }
                pix[tid] = z;
            }
            __syncthreads();
            if (li < K) {
                // The thread's lobe (axis mu, sharpness lam) is fixed, so the lobe-constant factors
                // come out of the pixel sum: with cd, cs the diffuse / specular chain factors,
                //   d/d(axis)      = lam * sum(cd w + cs q)
                //   d/d(sharpness) = mu . sum(cd w + cs q) - sum(A dg + B sg)
                //   d/d(amplitude) = sum(X dg) + sum(G sg)   (X, G carry 2.34 and sa)
                // (mu . w enters the sharpness term only through cd mu . w), and dg/dl =
                // il ((e1 + e2) - g) for g = (e1 - e2) il.  The e2 choice is uniform per wave.
                const SgLobe& L = lobes[li];
                const bool db = (cneed >> (2 * lw)) & 1u, sb = (cneed >> (2 * lw + 1)) & 1u;
                const int nit = ncovp / G;
                if (db && sb) sg_bwd_lobe_loop<true, true>(L, pix, grp, nit, G, accW, accC, accX, accG);
                else if (db) sg_bwd_lobe_loop<true, false>(L, pix, grp, nit, G, accW, accC, accX, accG);
                else if (sb) sg_bwd_lobe_loop<false, true>(L, pix, grp, nit, G, accW, accC, accX, accG);
                else sg_bwd_lobe_loop<false, false>(L, pix, grp, nit, G, accW, accC, accX, accG);
            }
            __syncthreads();
            cur = nxt;
        }
        if (want_lobes && li < K) {
            const SgLobe& L = lobes[li];
            // sum(cd w + cs q) = sum(cd p_d + cs p_s) + lam mu sum(cd + cs)
            accW[0] += L.gx * accC; accW[1] += L.gy * accC; accW[2] += L.gz * accC;
            acc[0] = accW[0] * L.lam;
            acc[1] = accW[1] * L.lam;
            acc[2] = accW[2] * L.lam;
            // sum(A dg + B sg) = a . sum(X dg) + a . sum(G sg): A and B are the lobe's amplitudes
            // dotted with the pixel's X and G, so the sharpness term's second sum comes from the
            // amplitude sums rather than from two more products per pixel
            float accS = 0.f;
#pragma unroll
            for (int k = 0; k < 4; k++) accS -= L.a[k] * (accX[k] + accG[k]);
            acc[3] = (L.mx * accW[0] + L.my * accW[1] + L.mz * accW[2]) + accS;
#pragma unroll
            for (int k = 0; k < 4; k++) acc[4 + k] = accX[k] + accG[k];
        }
        if (!want_lobes) return;
#pragma unroll
        for (int k = 0; k < SG_COLS; k++) red[tid][k] = acc[k];
        __syncthreads();
        for (int i = tid; i < K * SG_COLS; i += BS) {
            int s = i / SG_COLS, col = i % SG_COLS;   // s: the lobe's staged (sorted) position
            float sum = 0.f;
            for (int g2 = 0; g2 < G; g2++) sum += red[KP >= 4 ? (s / LPW) * 64 + g2 * LPW + s % LPW : g2 * KP + s][col];
            part[(size_t)vb.b.x * K * SG_COLS + (size_t)lorig[s] * SG_COLS + col] = col < 4 + C ? sum : 0.f;
        }
    }
};
__global__ void __launch_bounds__(BS) k_sg_bwd_lobes(size_t P, int C, int K, int KP, const float* rast, const float* axis, const float* sharp,
                                                     const float* amp, const float* normal, const float* diffuse, const float* sAxis,
                                                     const float* sSharp, const float* sAmp, const float* gout, const float* denv, float* gdiffuse,
                                                     float* part, int want_lobes, int gd_fresh) {
    KSgBwdLobesF{}(vb_hw(), P, C, K, KP, rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, gout, denv, gdiffuse, part, want_lobes, gd_fresh);
}

// ---------------------------------------------------------------- antialias
struct AAEdge { float alpha, ia, pax, pay, pbx, pby, sg; int va, vb, horiz, target_o, valid; };
constexpr int AA_MAXC = 4;

struct AAArgs {
    int W, H, depth, C, V;
    int v4;  // C == 4 and the per-pixel buffers 16-byte aligned: one float4 load per pixel
    float sx, sy;
    const float* proj;
    const u32* idx;
    const float* rast;
};

// The C (<= AA_MAXC) values of pixel q of a per-pixel buffer, zero-padded.
__device__ __forceinline__ void aa_load_px(const float* base, size_t q, int C, bool v4, float (&o)[AA_MAXC]) {
    if (v4) {
        const float4 t = ((const float4*)base)[q];
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    } else {
#pragma unroll
        for (int c = 0; c < AA_MAXC; c++) o[c] = c < C ? base[q * C + c] : 0.f;
    }
}

// The pixels of pair (current pixel (cx,cy), neighbour left/below): false when the neighbour is
// outside the image or both pixels hold the same triangle, else the front pixel pp, the other
// pixel op and the front pixel's rast record.
__device__ __forceinline__ bool aa_pair_pixels(const AAArgs& a, const float4* rastv, int cx, int cy, int vertical, int& pp, int& op,
                                               float4& front) {
    int pid = cy * a.W + cx;
    int nid;
    if (!vertical) { if (cx <= 0) return false; nid = pid - 1; } else { if (cy <= 0) return false; nid = pid - a.W; }
    float4 t = rastv[pid], nb = rastv[nid];
    if (nb.w == t.w) return false;
    pp = (int)t.w ? pid : nid;
    op = (int)nb.w ? pid : nid;
    if (pp == op) { if (t.z > nb.z) op = nid; else pp = nid; }
    front = pp == pid ? t : nb;
    return true;
}

// The edges of a pair whose pixels aa_pair_pixels found; returns the edge count written.
__device__ __forceinline__ int aa_pair_edges(const AAArgs& a, const float4* clipv, float4 front, int cx, int cy, int vertical, int pp, int op,
                                             AAEdge* E) {
#pragma clang fp contract(off)
    E[0].valid = E[1].valid = E[2].valid = 0;
    float fx, fy, ox, oy;
    if (!vertical) { fx = (float)cx + (pp > op ? .5f : -.5f); fy = (float)cy + .5f; ox = pp > op ? -1.f : 1.f; oy = 0.f; }
    else { fx = (float)cx + .5f; fy = (float)cy + (pp > op ? .5f : -.5f); ox = 0.f; oy = pp > op ? -1.f : 1.f; }
    int id = (int)front.w - 1;
    u32 vi[3] = {a.idx[3 * id], a.idx[3 * id + 1], a.idx[3 * id + 2]};
    float vx[3], vy[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float4 c = clipv[vi[k]];
        float iw = 1.f / c.w;
        vx[k] = (c.x * iw + 1.f) * a.sx - fx;
        vy[k] = (c.y * iw + 1.f) * a.sy - fy;
    }
    int ne = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int ia_ = k, ib_ = (k + 1) % 3;
        AAEdge& e = E[k];
        e.valid = 0;
        float pax = vx[ia_], pay = vy[ia_], pbx = vx[ib_], pby = vy[ib_];
        float av = pax * pby - pay * pbx;
        float oa = (pax - ox) * (pby - oy) - (pay - oy) * (pbx - ox);
        if (av * oa > 0.f) continue;
        float ex = pax - pbx, ey = pay - pby;
        float n = (ox + oy) * (ox - oy);
        if ((ex + ey) * (ex - ey) * n > 0.f) continue;
        float sg = ox + oy;
        ex *= sg; ey *= sg;
        float D = n > 0.f ? -ey : ex;
        float ia = (float)(1.0 / ((double)D + (D > 0 ? 1e-3 : -1e-3)));
        float alpha = av * ia - .5f;
        e.valid = 1;
        e.alpha = alpha; e.ia = ia; e.pax = pax; e.pay = pay; e.pbx = pbx; e.pby = pby; e.sg = sg;
        e.va = (int)vi[ia_]; e.vb = (int)vi[ib_]; e.horiz = n > 0.f; e.target_o = alpha > 0;
        ne++;
    }
    return ne;
}

// Tiled antialias: a workgroup owns a 16 x 16 pixel tile and evaluates every pixel pair that
// touches it once: the left and lower pair of each tile pixel (owned) plus the left pair of the
// column right of the tile and the lower pair of the row above it (halo, owned by the
// neighbouring tiles).  Each pair's contributions to its two pixels go to LDS; each pixel then
// adds the four pairs that touch it in the order (q, left), (q, below), (right, q), (above, q).
constexpr int AA_T = 16;
constexpr int AA_OWN = 2 * AA_T * AA_T;
constexpr int AA_PAIRS = AA_OWN + 2 * AA_T;
static_assert(AA_T * AA_T == BS, "one thread per tile pixel");

// Pair p of the tile with corner (x0, y0): its current pixel (the one whose left / lower
// neighbour forms the pair) and direction.
__device__ __forceinline__ void aa_tile_pair(int p, int x0, int y0, int& cx, int& cy, int& vertical) {
    if (p < AA_T * AA_T) { cx = x0 + (p % AA_T); cy = y0 + p / AA_T; vertical = 0; }
    else if (p < AA_OWN) { int r = p - AA_T * AA_T; cx = x0 + (r % AA_T); cy = y0 + r / AA_T; vertical = 1; }
    else if (p < AA_OWN + AA_T) { cx = x0 + AA_T; cy = y0 + (p - AA_OWN); vertical = 0; }
    else { cx = x0 + (p - AA_OWN - AA_T); cy = y0 + AA_T; vertical = 1; }
}

// LDS slot of the pair that touches tile pixel (tx, ty) as the k-th of its four pairs.
__device__ __forceinline__ int aa_pair_slot(int tx, int ty, int k) {
    switch (k) {
        case 0: return ty * AA_T + tx;
        case 1: return AA_T * AA_T + ty * AA_T + tx;
        case 2: return tx + 1 < AA_T ? ty * AA_T + tx + 1 : AA_OWN + ty;
        default: return ty + 1 < AA_T ? AA_T * AA_T + (ty + 1) * AA_T + tx : AA_OWN + AA_T + tx;
    }
}

struct KAaFwdF {
    static constexpr int kBS = BS, kMINW = 1;
    // target, gout, part (all or none): the L2 loss's grad out - target per pixel and the tile's
    // sum of squared grads (part[layer * tiles + tile]), the MSE pass fused into the antialias
    __device__ __forceinline__ void operator()(const VB& vb, AAArgs a, int tilesX, const float* in, float* out, const float* target, float* gout,
                                               float* part) const {
#pragma clang fp contract(off)
        __shared__ float4 to_cur[AA_PAIRS], to_nb[AA_PAIRS];
        __shared__ float wsq[BS / 64];
        const int s = vb.b.y;
        const unsigned tile = xcd_swizzle(vb.b.x, vb.g.x, (unsigned)tilesX);
        const int x0 = (tile % tilesX) * AA_T, y0 = (tile / tilesX) * AA_T;
        const float4* clipv = (const float4*)a.proj + (size_t)s * a.V;
        const float4* rastv = (const float4*)a.rast + (size_t)s * a.W * a.H;
        const float* inv = in + (size_t)s * a.W * a.H * a.C;
        const int C = a.C;
        const int tx = threadIdx.x % AA_T, ty = threadIdx.x / AA_T;
        const int px = x0 + tx, py = y0 + ty;
        const bool inside = px < a.W && py < a.H;
        const size_t q = inside ? (size_t)py * a.W + px : 0;
        // the own pixel's loss target, loaded first so that it arrives during the pair work
        float tg[AA_MAXC] = {0.f, 0.f, 0.f, 0.f};
        if (gout && inside) aa_load_px(target + (size_t)s * a.W * a.H * C, q, C, a.v4, tg);
        for (int p = threadIdx.x; p < AA_PAIRS; p += BS) {
            int cx, cy, vertical;
            aa_tile_pair(p, x0, y0, cx, cy, vertical);
            float tc[AA_MAXC] = {0.f, 0.f, 0.f, 0.f}, tn[AA_MAXC] = {0.f, 0.f, 0.f, 0.f};
            int pp = 0, op = 0;
            float4 front;
            AAEdge E[3];
            if (cx < a.W && cy < a.H && aa_pair_pixels(a, rastv, cx, cy, vertical, pp, op, front)) {
                // the two pixels' colours are loaded before the triangle's indices and vertices, so
                // the loads overlap that dependent chain
                float ip[AA_MAXC], io[AA_MAXC];
                aa_load_px(inv, (size_t)pp, C, a.v4, ip);
                aa_load_px(inv, (size_t)op, C, a.v4, io);
                aa_pair_edges(a, clipv, front, cx, cy, vertical, pp, op, E);
                int cur = cy * a.W + cx;
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    if (!E[e].valid) continue;
                    bool to_c = (E[e].target_o ? op : pp) == cur;
#pragma unroll
                    for (int c = 0; c < AA_MAXC; c++) {
                        if (c >= C) continue;
                        float v = E[e].alpha * (ip[c] - io[c]);
                        if (to_c) tc[c] += v; else tn[c] += v;
                    }
                }
            }
            to_cur[p] = make_float4(tc[0], tc[1], tc[2], tc[3]);
            to_nb[p] = make_float4(tn[0], tn[1], tn[2], tn[3]);
        }
        __syncthreads();
        float sq = 0.f;
        if (inside) {
            float acc[AA_MAXC];
            aa_load_px(inv, q, C, a.v4, acc);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float4 v = k < 2 ? to_cur[aa_pair_slot(tx, ty, k)] : to_nb[aa_pair_slot(tx, ty, k)];
                float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int c = 0; c < AA_MAXC; c++) acc[c] += vv[c];
            }
            float* o = out + ((size_t)s * a.W * a.H + q) * C;
#pragma unroll
            for (int c = 0; c < AA_MAXC; c++)
                if (c < C) o[c] = acc[c];
            if (gout) {
                float* g = gout + ((size_t)s * a.W * a.H + q) * C;
#pragma unroll
                for (int c = 0; c < AA_MAXC; c++)
                    if (c < C) {
                        const float d = acc[c] - tg[c];
                        g[c] = d;
                        sq += d * d;
                    }
            }
        }
        if (!part) return;
        sq = wave_sum(sq);
        if ((threadIdx.x & 63) == 0) wsq[threadIdx.x >> 6] = sq;
        __syncthreads();
        if (threadIdx.x == 0) part[(size_t)s * vb.g.x + tile] = ((wsq[0] + wsq[1]) + wsq[2]) + wsq[3];
    }
};
__global__ void __launch_bounds__(BS) k_aa_fwd(AAArgs a, int tilesX, const float* in, float* out, const float* target, float* gout, float* part) {
    KAaFwdF{}(vb_hw(), a, tilesX, in, out, target, gout, part);
}

// Backward of the tiled antialias.  Per pair, D = sum over its edges of alpha * gout[target]
// goes to the front pixel pp with + and to the other pixel op with -; the vertex (clip-space)
// grads of a pair's edges are added by the tile that owns the pair, summed per vertex in an LDS
// table first (neighbouring edge pixels share their edge's vertices) and flushed with four
// consecutive lanes per vertex (x, y, -, w).
constexpr int AA_VSLOTS = 256;

__device__ __forceinline__ void aa_vertex_grad(LdsTable<AA_VSLOTS, 3>& vt, float* gcs, int v, float gx, float gy, float gw) {
    int sl = lds_table_slot(vt, (u32)v);
    if (sl >= 0) {
        atomicAdd(&vt.val[0][sl], gx);
        atomicAdd(&vt.val[1][sl], gy);
        atomicAdd(&vt.val[2][sl], gw);
    } else {
        float* gp = gcs + 4 * (size_t)v;
        atomicAdd(gp, gx);
        atomicAdd(gp + 1, gy);
        atomicAdd(gp + 3, gw);
    }
}

struct KAaBwdF {
    static constexpr int kBS = BS, kMINW = 7;
    __device__ __forceinline__ void operator()(const VB& vb, AAArgs a, int tilesX, const float* in, const float* gout, float* gin, float* gproj) const {
#pragma clang fp contract(off)
        __shared__ float4 to_cur[AA_PAIRS];
        __shared__ LdsTable<AA_VSLOTS, 3> vt;
        __shared__ u32 vused[AA_VSLOTS], wave_count[BS / 64];
        const int s = vb.b.y;
        const unsigned tile = xcd_swizzle(vb.b.x, vb.g.x, (unsigned)tilesX);
        const int x0 = (tile % tilesX) * AA_T, y0 = (tile / tilesX) * AA_T;
        const float4* clipv = (const float4*)a.proj + (size_t)s * a.V;
        const float4* rastv = (const float4*)a.rast + (size_t)s * a.W * a.H;
        size_t off = (size_t)s * a.W * a.H * a.C;
        const float* inv = in + off;
        const float* gov = gout + off;
        float* gcs = gproj ? gproj + (size_t)s * a.V * 4 : nullptr;
        const int C = a.C;
        if (gcs) {
            lds_table_clear(vt);
            __syncthreads();
        }
        for (int p = threadIdx.x; p < AA_PAIRS; p += BS) {
            int cx, cy, vertical;
            aa_tile_pair(p, x0, y0, cx, cy, vertical);
            float D[AA_MAXC] = {0.f, 0.f, 0.f, 0.f};
            bool pp_cur = true;
            int pp = 0, op = 0;
            float4 front;
            AAEdge E[3];
            if (cx < a.W && cy < a.H && aa_pair_pixels(a, rastv, cx, cy, vertical, pp, op, front)) {
                pp_cur = pp == cy * a.W + cx;
                // both pixels' output grads (and colours, for the vertex grads) are loaded before the
                // triangle's indices and vertices, so the loads overlap that dependent chain
                const bool vgrad = gcs && p < AA_OWN;
                float gp_[AA_MAXC], go_[AA_MAXC], dc[AA_MAXC] = {0.f, 0.f, 0.f, 0.f};
                aa_load_px(gov, (size_t)pp, C, a.v4, gp_);
                aa_load_px(gov, (size_t)op, C, a.v4, go_);
                if (vgrad) {
                    float ip[AA_MAXC], io[AA_MAXC];
                    aa_load_px(inv, (size_t)pp, C, a.v4, ip);
                    aa_load_px(inv, (size_t)op, C, a.v4, io);
#pragma unroll
                    for (int c = 0; c < AA_MAXC; c++) dc[c] = ip[c] - io[c];
                }
                aa_pair_edges(a, clipv, front, cx, cy, vertical, pp, op, E);
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    if (!E[e].valid) continue;
                    const bool to_o = E[e].target_o;
                    float al = E[e].alpha;
#pragma unroll
                    for (int c = 0; c < AA_MAXC; c++)
                        if (c < C) D[c] += al * (to_o ? go_[c] : gp_[c]);
                    if (vgrad) {
                        float d = 0.f;
#pragma unroll
                        for (int c = 0; c < AA_MAXC; c++)
                            if (c < C) d += (to_o ? go_[c] : gp_[c]) * dc[c];
                        const AAEdge& ed = E[e];
                        float ia = ed.ia, av = ed.pax * ed.pby - ed.pay * ed.pbx;
                        float dDax, dDay, dDbx, dDby;
                        if (ed.horiz) { dDax = 0.f; dDay = -ed.sg; dDbx = 0.f; dDby = ed.sg; }
                        else { dDax = ed.sg; dDay = 0.f; dDbx = -ed.sg; dDby = 0.f; }
                        float k2 = av * ia * ia;
                        float gax = d * (ia * ed.pby - k2 * dDax), gay = d * (-ia * ed.pbx - k2 * dDay);
                        float gbx = d * (-ia * ed.pay - k2 * dDbx), gby = d * (ia * ed.pax - k2 * dDby);
                        int vv[2] = {ed.va, ed.vb};
                        float gxs[2] = {gax, gbx}, gys[2] = {gay, gby};
                        for (int j = 0; j < 2; j++) {
                            float4 c = clipv[vv[j]];
                            float iw = 1.f / c.w;
                            float gvx = gxs[j] * a.sx * iw, gvy = gys[j] * a.sy * iw;
                            aa_vertex_grad(vt, gcs, vv[j], gvx, gvy, -(gvx * c.x + gvy * c.y) * iw);
                        }
                    }
                }
            }
            // contribution to the current pixel; the neighbour receives its negation
            to_cur[p] = pp_cur ? make_float4(D[0], D[1], D[2], D[3]) : make_float4(-D[0], -D[1], -D[2], -D[3]);
        }
        __syncthreads();
        const int tx = threadIdx.x % AA_T, ty = threadIdx.x / AA_T;
        const int px = x0 + tx, py = y0 + ty;
        if (px < a.W && py < a.H) {
            size_t q = (size_t)py * a.W + px;
            float acc[AA_MAXC];
            aa_load_px(gov, q, C, a.v4, acc);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float4 v = to_cur[aa_pair_slot(tx, ty, k)];
                float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int c = 0; c < AA_MAXC; c++) acc[c] = k < 2 ? acc[c] + vv[c] : acc[c] - vv[c];
            }
            float* o = gin + off + q * C;
#pragma unroll
            for (int c = 0; c < AA_MAXC; c++)
                if (c < C) o[c] = acc[c];
        }
        // the pair loop's LDS table updates finished at the barrier above
        if (gcs) lds_vertex_grad_flush(vt, vused, wave_count, gcs);
    }
};
__global__ void __launch_bounds__(BS, 7) k_aa_bwd(AAArgs a, int tilesX, const float* in, const float* gout, float* gin, float* gproj) { KAaBwdF{}(vb_hw(), a, tilesX, in, gout, gin, gproj); }

// ---------------------------------------------------------------- MSE
// grad = pred - target: four elements per thread (one 16-byte access each when `vec`) and a
// partial sum of grad^2 per block.  (A last-block-finishes reduction would need one device
// atomic per block on a single counter; those serialise, so the partials are summed by a
// second launch.)
struct KMseF {
    static constexpr int kBS = BS, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, size_t n, const float* pred, const float* target, float* grad, float* part, int vec) const {
        const size_t t = (size_t)vb.b.x * BS + threadIdx.x;
        float s = 0.f;
        if (vec) {
            if (4 * t < n) {
                const float4 a = ((const float4*)pred)[t], b = ((const float4*)target)[t];
                const float4 g = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
                ((float4*)grad)[t] = g;
                s = ((g.x * g.x + g.y * g.y) + g.z * g.z) + g.w * g.w;
            }
        } else {
            for (int k = 0; k < 4; k++) {
                const size_t i = 4 * t + k;
                if (i < n) { float g = pred[i] - target[i]; grad[i] = g; s += g * g; }
            }
        }
        s = wave_sum(s);
        __shared__ float sh[BS / 64];
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) part[vb.b.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    }
};
__global__ void __launch_bounds__(BS) k_mse(size_t n, const float* pred, const float* target, float* grad, float* part, int vec) { KMseF{}(vb_hw(), n, pred, target, grad, part, vec); }

// loss = 0.5 * the sum of the nb block partials, in a fixed order (one 1024-thread workgroup;
// thread t sums its contiguous share with four interleaved accumulators so the loads overlap)
struct KMseFinalF {
    static constexpr int kBS = 1024, kMINW = 1;
    __device__ __forceinline__ void operator()(const VB& vb, const float* part, size_t nb, float* loss) const { mse_final_block<1024>(part, nb, loss); }
};
__global__ void __launch_bounds__(1024) k_mse_final(const float* part, size_t nb, float* loss) { KMseFinalF{}(vb_hw(), part, nb, loss); }

}  // namespace

// =============================================================================== C ABI
extern "C" const char* vcl_status_string(int s) {
    switch (s) {
        case VCL_OK: return "ok";
        case VCL_ERR_INVALID: return "invalid argument";
        case VCL_ERR_OOM: return "out of memory / workspace too small";
        case VCL_ERR_HIP: return "HIP runtime error";
        case VCL_ERR_OVERFLOW: return "device capacity overflow";
        default: return "unknown status";
    }
}

extern "C" const char* vcl_build_info(void) { return "vcl_amd 0.1 gfx950"; }

int vcl_project_forward_dev(const vcl_project_params* p, float* zero, double* bt, double rhom, double rhov, const WsClear* rc,
                            vcl_stream_t stream) {
    VCL_REQUIRE(p && p->vec && p->mat && p->out && p->vecNum >= 0 && (p->dimention == 3 || p->dimention == 4));
    int views = p->views < 1 ? 1 : p->views;
    size_t n = (size_t)p->vecNum * views;
    WsClear c;
    std::memset(&c, 0, sizeof(c));
    if (rc) c = *rc;
    const size_t total = std::max(n, c.n());
    if (!total && !bt) return VCL_OK;
    vlaunch<KProjectFwdF>(k_project_fwd, dim3(total ? cdiv(total, BS) : 1), dim3(BS), 0, as_hip(stream), p->vecNum, p->dimention, views, p->vec, p->mat,
                       p->out, zero, bt, rhom, rhov, c);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_project_forward(const vcl_project_params* p, vcl_stream_t stream) {
    return vcl_project_forward_dev(p, nullptr, nullptr, 1.0, 1.0, nullptr, stream);
}

extern "C" int vcl_project_backward(const vcl_project_params* p, const vcl_project_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->vec && p->mat && g->out && (p->dimention == 3 || p->dimention == 4));
    int views = p->views < 1 ? 1 : p->views;
    if (p->vecNum <= 0 || (!g->vec && !g->mat)) return VCL_OK;
    vlaunch<KProjectBwdF>(k_project_bwd, dim3(cdiv(p->vecNum, BS)), dim3(BS), 0, as_hip(stream), p->vecNum, p->dimention, views, p->vec, p->mat, g->out,
                       g->vec, g->mat);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_interpolate_forward(const vcl_interpolate_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->rast && p->attr && p->idx && p->out && p->dimention > 0);
    VCL_REQUIRE(!p->enableDA || (p->rastDB && p->outDA));
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    vlaunch<KInterpFwdF>(k_interp_fwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->dimention, p->enableDA, p->rast, p->rastDB, p->attr,
                       p->idx, p->out, p->outDA);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_interpolate_backward(const vcl_interpolate_params* p, const vcl_interpolate_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->rast && p->attr && p->idx && g->out);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P || (!g->attr && !g->rast)) return VCL_OK;
    vlaunch<KInterpBwdF>(k_interp_bwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->dimention, p->rast, p->attr, p->idx, g->out, g->attr,
                       g->rast);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

static int tex_args(const vcl_texturemap_params* p, TexArgs& t) {
    VCL_REQUIRE(p->texwidth > 0 && p->texheight > 0 && p->channel > 0 && p->miplevel >= 1 && p->miplevel <= VCL_TEX_MAX_MIP_LEVEL);
    t.tw = p->texwidth; t.th = p->texheight; t.C = p->channel; t.L = p->miplevel;
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) t.tex[l] = l < t.L ? p->texture[l] : nullptr;
    for (int l = 0; l < t.L; l++) VCL_REQUIRE(t.tex[l]);
    if (t.L > 1) VCL_REQUIRE(p->uvDA);
    return VCL_OK;
}

extern "C" int vcl_texturemap_forward(const vcl_texturemap_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->rast && p->uv && p->out);
    TexArgs t;
    int st = tex_args(p, t);
    if (st) return st;
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    if (t.C == 4)
        vlaunch<KTexFwd4F>(k_tex_fwd4, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, t, p->rast, p->uv, p->uvDA, p->out);
    else
        vlaunch<KTexFwdF>(k_tex_fwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, t, p->rast, p->uv, p->uvDA, p->out);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

int vcl_interp_tex_dev(const vcl_interpolate_params* ip, const vcl_texturemap_params* tp, float* zero, vcl_stream_t stream) {
    VCL_REQUIRE(ip && tp && ip->rast && ip->rastDB && ip->attr && ip->idx && ip->out && ip->outDA && tp->out);
    VCL_REQUIRE(ip->dimention == 2 && ip->enableDA && tp->channel == 4 && tp->uv == ip->out && tp->uvDA == ip->outDA && tp->rast == ip->rast);
    VCL_REQUIRE(ip->width == tp->width && ip->height == tp->height && ip->depth == tp->depth);
    TexArgs t;
    int st = tex_args(tp, t);
    if (st) return st;
    size_t P = (size_t)ip->width * ip->height * ip->depth;
    if (!P) return VCL_OK;
    vlaunch<KInterpTex4F<INTERP_TEX_NPX>>(k_interp_tex4, dim3(cdiv(P, (size_t)BS * INTERP_TEX_NPX)), dim3(BS), 0, as_hip(stream), P, ip->rast, ip->rastDB, ip->attr, ip->idx, ip->out, ip->outDA,
                       t, tp->out, zero);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// ib.grast not null: the interpolate backward to run inside the tiled kernel.  *fused_done tells
// whether it ran; when the tiled kernel does not apply it stays false and the caller runs the
// interpolate backward itself.
// stage (may be null): a segment stage to run in the tiled kernel's launch, or on its own otherwise.
static int texturemap_backward(const vcl_texturemap_params* p, const vcl_texturemap_grad_params* g, const InterpBwdArgs& ib, bool* fused_done,
                               vcl_stream_t stream, const VclColsumStage* stage = nullptr) {
    if (fused_done) *fused_done = false;
    VCL_REQUIRE(p && g && p->rast && p->uv && g->out);
    TexArgs t;
    int st = tex_args(p, t);
    if (st) return st;
    TexGradArgs ga;
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) ga.tex[l] = l < t.L ? g->texture[l] : nullptr;
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    bool all_or_none = true;
    for (int l = 1; l < t.L; l++) all_or_none &= ((ga.tex[l] != nullptr) == (ga.tex[0] != nullptr));
    if (t.C == 4 && (size_t)t.tw * t.th <= (1u << 24) && all_or_none) {
        const int rows = g_vcl_tuning.splat_rows;
        int tilesX = (p->width + 15) / 16, tilesY = (p->height + rows - 1) / rows;
        dim3 grid(tilesX * tilesY, p->depth);
        float* gda = t.L > 1 ? g->uvDA : nullptr;
        // fused: the uv grad is consumed in the kernel by the interpolate backward, not stored
        float* guv = ib.grast ? nullptr : g->uv;
        if (stage && rows == 16) {
            const int nseg = colsum_nseg(*stage);
            const size_t nb = (size_t)tilesX * tilesY * p->depth + cdiv(stage->cols, 64) * (size_t)nseg;
            vlaunch<KTexBwdStageF<16>>(k_tex_bwd_stage<16>, dim3((unsigned)nb), dim3(BS), 0, as_hip(stream), tilesX * tilesY, p->depth, p->width,
                                       p->height, tilesX, t, ga, p->rast, p->uv, p->uvDA, g->out, guv, gda, g->daMode, ib, *stage, nseg);
            stage = nullptr;
        } else if (rows == 64)
            vlaunch<KTexBwd4TiledF<64>>(k_tex_bwd4_tiled<64>, grid, dim3(BS), 0, as_hip(stream), p->width, p->height, tilesX, t, ga, p->rast, p->uv, p->uvDA,
                               g->out, guv, gda, g->daMode, ib);
        else if (rows == 32)
            vlaunch<KTexBwd4TiledF<32>>(k_tex_bwd4_tiled<32>, grid, dim3(BS), 0, as_hip(stream), p->width, p->height, tilesX, t, ga, p->rast, p->uv, p->uvDA,
                               g->out, guv, gda, g->daMode, ib);
        else
            vlaunch<KTexBwd4TiledF<16>>(k_tex_bwd4_tiled<16>, grid, dim3(BS), 0, as_hip(stream), p->width, p->height, tilesX, t, ga, p->rast, p->uv, p->uvDA,
                               g->out, guv, gda, g->daMode, ib);
        if (fused_done) *fused_done = ib.grast != nullptr;
    } else {
        vlaunch<KTexBwdF>(k_tex_bwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, t, ga, p->rast, p->uv, p->uvDA, g->out, g->uv,
                           t.L > 1 ? g->uvDA : nullptr, g->daMode);
    }
    VCL_LAUNCH_CHECK();
    if (stage && colsum_launch(*stage, as_hip(stream))) return VCL_ERR_HIP;
    return VCL_OK;
}

extern "C" int vcl_texturemap_backward(const vcl_texturemap_params* p, const vcl_texturemap_grad_params* g, vcl_stream_t stream) {
    return texturemap_backward(p, g, InterpBwdArgs{nullptr, nullptr, nullptr, nullptr}, nullptr, stream);
}

int vcl_texture_interp_backward_dev(const vcl_texturemap_params* tp, const vcl_texturemap_grad_params* tg, const vcl_interpolate_params* ip,
                                    const vcl_interpolate_grad_params* ig, vcl_stream_t stream, const vcl_sg_lobe_sums* stage) {
    VCL_REQUIRE(tp && tg && ip && ig && ip->rast && ip->attr && ip->idx && ig->out);
    const VclColsumStage* cs = stage && stage->stage_valid ? &stage->stage : nullptr;
    // fusable: the interpolate backward's output grad is this texture backward's uv grad, over
    // the same pixels, with a 2-channel attribute and only the rast grad wanted
    const bool fusable = ip->dimention == 2 && ig->out == tg->uv && !ig->attr && ig->rast && ip->rast == tp->rast && ip->width == tp->width &&
                         ip->height == tp->height && ip->depth == tp->depth;
    bool done = false;
    int st = texturemap_backward(tp, tg, fusable ? InterpBwdArgs{ip->rast, ip->idx, ip->attr, ig->rast} : InterpBwdArgs{nullptr, nullptr, nullptr, nullptr},
                                 &done, stream, cs);
    if (st || done) return st;
    return vcl_interpolate_backward(ip, ig, stream);
}

extern "C" int vcl_texture_mip_levels(int w, int h, int requested) {
    unsigned v = (unsigned)(w | h);
    int lsb = 0;
    while (lsb < 31 && !((v >> lsb) & 1u)) lsb++;
    int maxl = lsb + 1 < 16 ? lsb + 1 : 16;
    return requested < 1 ? 1 : (maxl < requested ? maxl : requested);
}

extern "C" int vcl_texture_build_mip(const vcl_texture* t, vcl_stream_t stream) {
    VCL_REQUIRE(t && t->miplevel >= 1 && t->miplevel <= VCL_TEX_MAX_MIP_LEVEL && t->channel > 0);
    for (int l = 1; l < t->miplevel; l++) {
        VCL_REQUIRE(t->texture[l] && t->texture[l - 1]);
        int w = t->width >> l, h = t->height >> l;
        size_t n = (size_t)w * h * t->channel;
        if (!n) continue;
        vlaunch<KDownsampleF>(k_downsample, dim3(cdiv(n, BS)), dim3(BS), 0, as_hip(stream), t->texture[l - 1], t->texture[l], w, h, t->channel,
                           (float*)nullptr);
        VCL_LAUNCH_CHECK();
    }
    return VCL_OK;
}

extern "C" int vcl_texture_clamp(const vcl_texture* t, float lo, float hi, vcl_stream_t stream) {
    VCL_REQUIRE(t && t->texture[0]);
    size_t n = (size_t)t->width * t->height * t->channel;
    if (n) {
        vlaunch<KClampF>(k_clamp, dim3(cdiv(n, BS)), dim3(BS), 0, as_hip(stream), t->texture[0], n, lo, hi);
        VCL_LAUNCH_CHECK();
    }
    return vcl_texture_build_mip(t, stream);
}

extern "C" int vcl_texture_grad_sumup(const vcl_texture* g, int quarter, vcl_stream_t stream) {
    VCL_REQUIRE(g && g->miplevel >= 1 && g->miplevel <= VCL_TEX_MAX_MIP_LEVEL);
    for (int l = g->miplevel - 1; l > 0; l--) {
        VCL_REQUIRE(g->texture[l] && g->texture[l - 1]);
        int w = g->width >> l, h = g->height >> l;
        size_t n = (size_t)w * h * g->channel;
        if (!n) continue;
        vlaunch<KFoldF>(k_fold, dim3(cdiv(n, BS)), dim3(BS), 0, as_hip(stream), g->texture[l], g->texture[l - 1], w, h, g->channel,
                           quarter ? 0.25f : 1.f);
        VCL_LAUNCH_CHECK();
    }
    return VCL_OK;
}

extern "C" int vcl_texture_grad_clear(const vcl_texture* g, vcl_stream_t stream) {
    VCL_REQUIRE(g && g->miplevel >= 1 && g->miplevel <= VCL_TEX_MAX_MIP_LEVEL);
    for (int l = 0; l < g->miplevel; l++) {
        size_t n = (size_t)(g->width >> l) * (g->height >> l) * g->channel;
        int st = vcl_fill_async(g->texture[l], n * sizeof(float), as_hip(stream));
        if (st) return st;
    }
    return VCL_OK;
}

// bt_dev: device {rhomt, rhovt} used instead of the per-tensor host values (null: host values).
extern "C" int vcl_adam_step_multi_dev(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, const double* bt_dev,
                                       int zero_grad, vcl_stream_t stream) {
    return vcl_adam_step_multi_proj_dev(n, p, h, bt_dev, zero_grad, nullptr, nullptr, nullptr, stream, -1, nullptr);
}

int vcl_adam_step_multi_proj_dev(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, const double* bt_dev, int zero_grad,
                                 const vcl_project_params* proj, const vcl_loss_params* loss, const vcl_sg_lobe_sums* lobes,
                                 vcl_stream_t stream, long long loss_parts, const VclMipTail* mip_tail) {
    if (loss) VCL_REQUIRE(loss->buffer && loss->loss && loss->size >= 0 && n <= ADAM_MAX);
    if (lobes) {
        // tensors 1..3: the lobe axis (K x 3), sharpness (K) and amplitude (K x C) grads
        VCL_REQUIRE(lobes->seg && lobes->nseg >= 0 && lobes->C >= 1 && lobes->C <= 4 && n >= 4 && n <= ADAM_MAX && p[1].size == 3LL * lobes->K &&
                    p[2].size == (long long)lobes->K && p[3].size == (long long)lobes->K * lobes->C);
    }
    VCL_REQUIRE(n >= 0 && (n == 0 || (p && h)));
    if (proj) {
        // tensor 0 must be the projected positions, 3 floats per vertex, with 4-float clip grads
        VCL_REQUIRE(n >= 1 && proj->dimention == 4 && proj->vec == p[0].param && p[0].size == 3 * (long long)proj->vecNum && proj->mat &&
                    proj->out);
    }
    for (int base = 0; base < n; base += ADAM_MAX) {
        AdamArgs a;
        a.bt = bt_dev;
        a.zero_grad = zero_grad;
        a.pgclip = base == 0 && proj ? proj->out : nullptr;
        a.pmats = proj ? proj->mat : nullptr;
        a.pviews = proj ? (proj->views < 1 ? 1 : proj->views) : 0;
        a.pV = proj ? proj->vecNum : 0;
        a.loss = base == 0 && loss ? loss->loss : nullptr;
        a.loss_part = loss ? loss->buffer : nullptr;
        a.loss_nb = !loss ? 0 : loss_parts >= 0 ? (size_t)loss_parts : loss->size > 0 ? cdiv((size_t)loss->size, (size_t)BS * 4) : 0;
        a.lseg = base == 0 && lobes ? lobes->seg : nullptr;
        a.lnseg = lobes ? lobes->nseg : 0;
        a.lK = lobes ? lobes->K : 0;
        a.lC = lobes ? lobes->C : 1;
        a.n = n - base < ADAM_MAX ? n - base : ADAM_MAX;
        long long blocks = 0;
        for (int i = 0; i < a.n; i++) {
            const vcl_optimize_params& q = p[base + i];
            const vcl_adam_hyper_params& hh = h[base + i];
            VCL_REQUIRE(q.size >= 0 && (q.size == 0 || (q.param && q.grad && hh.m && hh.v)));
            a.param[i] = q.param; a.grad[i] = q.grad; a.m[i] = hh.m; a.v[i] = hh.v; a.size[i] = q.size;
            a.rhom[i] = hh.rhom; a.rhov[i] = hh.rhov; a.rhomt[i] = hh.rhomt; a.rhovt[i] = hh.rhovt; a.eta[i] = hh.eta; a.eps[i] = hh.eps;
            a.block_start[i] = blocks;
            a.per_thread[i] = base == 0 && lobes && i >= 1 && i <= 3 ? 1 : ADAM_PER_BLOCK / BS;
            blocks += (q.size + (long long)BS * a.per_thread[i] - 1) / ((long long)BS * a.per_thread[i]);
        }
        a.block_start[a.n] = blocks;
        const bool tail = base == 0 && mip_tail && mip_tail->valid;
        if (tail) {
            MipArgs m;
            m.src = (const float4*)mip_tail->src;
            m.sw = mip_tail->sw; m.sh = mip_tail->sh; m.nl = mip_tail->nl;
            for (int k = 0; k < 6; k++) { m.dst[k] = (float4*)mip_tail->dst[k]; m.zg[k] = (float4*)mip_tail->zg[k]; }
            vlaunch<KAdamMipF>(k_adam_mip, dim3((unsigned)(blocks + (a.loss ? 1 : 0) + 1)), dim3(BS), 0, as_hip(stream), a, m);
            VCL_LAUNCH_CHECK();
        } else if (blocks > 0 || a.loss) {
            vlaunch<KAdamF>(k_adam, dim3((unsigned)(blocks + (a.loss ? 1 : 0))), dim3(BS), 0, as_hip(stream), a);
            VCL_LAUNCH_CHECK();
        }
    }
    return VCL_OK;
}

extern "C" int vcl_adam_step_multi(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, vcl_stream_t stream) {
    return vcl_adam_step_multi_dev(n, p, h, nullptr, 0, stream);
}

extern "C" int vcl_adam_step(const vcl_optimize_params* p, const vcl_adam_hyper_params* h, vcl_stream_t stream) {
    return vcl_adam_step_multi(1, p, h, stream);
}

// part bit 1: fold + Adam + clamp on level 0; part bit 2: rebuild mips (and clear coarse grads).
int vcl_texture_adam_update_parts(const vcl_texture* tex, const vcl_texture* grad, const vcl_adam_hyper_params* h, float lo, float hi,
                                  int quarter, int part, const double* bt_dev, vcl_stream_t stream, VclMipTail* tail) {
    if (tail) tail->valid = 0;
    VCL_REQUIRE(tex && grad && h && h->m && h->v && tex->miplevel == grad->miplevel && tex->width == grad->width && tex->height == grad->height &&
                tex->channel == grad->channel && tex->miplevel >= 1 && tex->miplevel <= VCL_TEX_MAX_MIP_LEVEL);
    TexUpdArgs a;
    a.tw = tex->width; a.th = tex->height; a.C = tex->channel; a.L = tex->miplevel;
    a.tex0 = tex->texture[0];
    a.tex1 = a.L > 1 ? tex->texture[1] : nullptr;
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) a.grad[l] = l < a.L ? grad->texture[l] : nullptr;
    for (int l = 0; l < a.L; l++) VCL_REQUIRE(a.grad[l] && tex->texture[l]);
    a.m = h->m; a.v = h->v;
    a.rhom = h->rhom; a.rhov = h->rhov; a.rhomt = h->rhomt; a.rhovt = h->rhovt; a.eta = h->eta; a.eps = h->eps;
    a.lo = lo; a.hi = hi; a.scale = quarter ? 0.25f : 1.f;
    a.bt = bt_dev;
    size_t n = (size_t)a.tw * a.th * a.C;
    hipStream_t s = as_hip(stream);
    bool vec4 = a.C == 4;
    for (int l = 0; l < a.L; l++) vec4 &= ((uintptr_t)a.grad[l] % 16 == 0) && ((uintptr_t)tex->texture[l] % 16 == 0);
    vec4 &= ((uintptr_t)a.m % 16 == 0) && ((uintptr_t)a.v % 16 == 0);
    // part & 4: level 1 is built by the C == 4 update (k_tex_update4m) instead of the mip pass;
    // both calls of a split update must pass it
    const bool quad = vec4 && a.L > 1 && (part & 4) && a.tw % 2 == 0 && a.th % 2 == 0 && a.th / 2 <= 65535;
    if (part & 1) {
        if (quad)
            vlaunch<KTexUpdate4mF>(k_tex_update4m, dim3((unsigned)cdiv((size_t)a.tw, 128), (unsigned)(a.th / 2)), dim3(BS), 0, s, a);
        else if (vec4)
            vlaunch<KTexUpdate4F>(k_tex_update4, dim3(cdiv((size_t)a.tw * a.th, BS)), dim3(BS), 0, s, a);
        else
            vlaunch<KTexUpdateF>(k_tex_update, dim3(cdiv(n, BS)), dim3(BS), 0, s, a);
        VCL_LAUNCH_CHECK();
    }
    // rebuild mips from the new level 0; the same pass zeroes the coarse grads.  C == 4: level 1
    // with one thread per texel, then up to five or six levels per launch (k_mip_fused)
    if (!(part & 2)) return VCL_OK;
    if (vec4) {
        int l = quad ? 2 : 1;
        if (l == 1 && l < a.L) {
            int w = a.tw >> 1, hh = a.th >> 1;
            vlaunch<KDownsample4F>(k_downsample4, dim3(cdiv((size_t)w * hh, BS)), dim3(BS), 0, s, (const float4*)tex->texture[0],
                               (float4*)tex->texture[1], w, hh, (float4*)a.grad[1]);
            VCL_LAUNCH_CHECK();
            l = 2;
        }
        while (l < a.L) {
            MipArgs m;
            m.src = (const float4*)tex->texture[l - 1];
            m.sw = a.tw >> (l - 1);
            m.sh = a.th >> (l - 1);
            // a source that one 64 x 64 block covers in one pass (six levels), any other in 32 x 32
            // blocks (five levels per pass), so the launch sequence does not depend on the size
            const bool b32 = !(m.sw <= 64 && m.sh <= 64);
            m.nl = std::min(b32 ? 5 : 6, a.L - l);
            for (int k = 0; k < 6; k++) {
                m.dst[k] = k < m.nl ? (float4*)tex->texture[l + k] : nullptr;
                m.zg[k] = k < m.nl ? (float4*)a.grad[l + k] : nullptr;
            }
            if (tail && l + m.nl >= a.L && m.sw <= 64 && m.sh <= 64) {
                // the last pass is one workgroup: the caller runs it (in its Adam launch)
                tail->valid = 1;
                tail->src = (const float*)m.src; tail->sw = m.sw; tail->sh = m.sh; tail->nl = m.nl;
                for (int k = 0; k < 6; k++) { tail->dst[k] = (float*)m.dst[k]; tail->zg[k] = (float*)m.zg[k]; }
                break;
            }
            if (b32)
                vlaunch<KMipFusedF<32>>(k_mip_fused<32>, dim3(cdiv(m.sw, 32), cdiv(m.sh, 32)), dim3(BS), 0, s, m);
            else
                vlaunch<KMipFusedF<64>>(k_mip_fused<64>, dim3(cdiv(m.sw, 64), cdiv(m.sh, 64)), dim3(BS), 0, s, m);
            VCL_LAUNCH_CHECK();
            l += m.nl;
        }
        return VCL_OK;
    }
    for (int l = 1; l < a.L; l++) {
        int w = a.tw >> l, hh = a.th >> l;
        size_t m = (size_t)w * hh * a.C;
        if (!m) continue;
        vlaunch<KDownsampleF>(k_downsample, dim3(cdiv(m, BS)), dim3(BS), 0, s, tex->texture[l - 1], tex->texture[l], w, hh, a.C, a.grad[l]);
        VCL_LAUNCH_CHECK();
    }
    return VCL_OK;
}

extern "C" int vcl_texture_adam_update(const vcl_texture* tex, const vcl_texture* grad, const vcl_adam_hyper_params* h, float lo, float hi,
                                       int quarter, vcl_stream_t stream) {
    return vcl_texture_adam_update_parts(tex, grad, h, lo, hi, quarter, 3 | 4, nullptr, stream, nullptr);
}

extern "C" int vcl_normal_axis_forward(const vcl_normal_axis_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->rast && p->rot && p->normal && p->normalidx && p->out);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    vlaunch<KNormalAxisF>(k_normal_axis, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->rast, p->rot, p->normal, p->normalidx, p->out);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// inverse(mat3(P) * mat3(V)), column-major, computed in double and rounded (glm's float
// inverse is not reproduced bit for bit).
static void pv_inverse(const float* Pm, const float* Vm, float* out9) {
    double A[3][3];  // math matrix A[row][col] = sum_k P[row][k] V[k][col]
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += (double)Pm[k * 4 + r] * (double)Vm[c * 4 + k];
            A[r][c] = s;
        }
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                 A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    double id = 1.0 / det;
    double I[3][3];
    I[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
    I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    I[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
    I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    I[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
    I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) out9[c * 3 + r] = (float)I[r][c];
}

extern "C" int vcl_view_axis_forward(const vcl_view_axis_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->out && p->view && p->projection && p->depth >= 1);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    VCL_REQUIRE(p->depth <= 8);
    Mat3x8 m;
    for (int s = 0; s < p->depth; s++) pv_inverse(p->projection + 16 * s, p->view + 16 * s, m.m[s]);
    vlaunch<KViewAxisF>(k_view_axis, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), p->width, p->height, p->depth, (const float*)nullptr, m,
                       p->out);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// Device-side pvinv variant used by the fit engine (one 3x3 per layer already on the device).
extern "C" int vcl_view_axis_forward_dev(int W, int H, int depth, const float* pvinv_dev, float* out, vcl_stream_t stream) {
    VCL_REQUIRE(pvinv_dev && out);
    size_t P = (size_t)W * H * depth;
    if (!P) return VCL_OK;
    Mat3x8 m;
    vlaunch<KViewAxisF>(k_view_axis, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), W, H, depth, pvinv_dev, m, out);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

int vcl_shade_prep_dev(const vcl_normal_axis_params* np, const float* pvinv_dev, const vcl_sg_specular_params* sp, vcl_stream_t stream) {
    VCL_REQUIRE(np && sp && pvinv_dev && np->rast && np->rot && np->normal && np->normalidx && np->out && sp->view && sp->roughness &&
                sp->outAxis && sp->outSharpness && sp->outAmplitude);
    VCL_REQUIRE(np->width == sp->width && np->height == sp->height && np->depth == sp->depth && sp->normal == np->out && sp->rast == np->rast);
    size_t P = (size_t)np->width * np->height * np->depth;
    if (!P) return VCL_OK;
    vlaunch<KShadePrepF>(k_shade_prep, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), np->width, np->height, np->depth, np->rast, np->rot, np->normal,
                       np->normalidx, pvinv_dev, sp->roughness, sp->ior, np->out, const_cast<float*>(sp->view), sp->outAxis, sp->outSharpness,
                       sp->outAmplitude);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

int vcl_pixel_prep_dev(const vcl_rasterize_params* rp, const unsigned long long* keys, const vcl_interpolate_params* ip,
                       const vcl_texturemap_params* tp, const vcl_normal_axis_params* np, const float* pvinv_dev, const vcl_sg_specular_params* sp,
                       vcl_stream_t stream) {
    VCL_REQUIRE(rp && keys && ip && tp && np && sp && pvinv_dev);
    VCL_REQUIRE(rp->out && rp->outDB && rp->enableDB && rp->proj && rp->idx && ip->rast == rp->out && ip->rastDB == rp->outDB);
    VCL_REQUIRE(rp->width == ip->width && rp->height == ip->height && rp->depth == ip->depth && rp->idx == ip->idx);
    VCL_REQUIRE(ip->attr && ip->idx && ip->out && ip->outDA && tp->out);
    VCL_REQUIRE(ip->dimention == 2 && ip->enableDA && tp->channel == 4 && tp->uv == ip->out && tp->uvDA == ip->outDA && tp->rast == ip->rast);
    VCL_REQUIRE(np->rot && np->normal && np->normalidx == ip->idx && np->out && sp->roughness && sp->outAxis && sp->outSharpness && sp->outAmplitude);
    VCL_REQUIRE(np->rast == ip->rast && sp->rast == ip->rast && sp->normal == np->out);
    VCL_REQUIRE(ip->width == tp->width && ip->height == tp->height && ip->depth == tp->depth && np->width == ip->width && np->height == ip->height &&
                np->depth == ip->depth && sp->width == ip->width && sp->height == ip->height && sp->depth == ip->depth);
    TexArgs t;
    int st = tex_args(tp, t);
    if (st) return st;
    size_t P = (size_t)ip->width * ip->height * ip->depth;
    if (!P) return VCL_OK;
    const ResolveArgs ra = {keys, rp->proj, rp->projNum, rp->xs, rp->ys, rp->out, rp->outDB};
    vlaunch<KPixelPrep4F<PIXEL_PREP_NPX>>(k_pixel_prep4, dim3(cdiv(P, (size_t)BS * PIXEL_PREP_NPX)), dim3(BS), 0, as_hip(stream), ip->width,
                                          ip->height, ip->depth, ra, ip->attr, ip->idx, ip->out, ip->outDA, t, tp->out, np->rot, np->normal,
                                          pvinv_dev, sp->roughness, sp->ior, np->out, const_cast<float*>(sp->view), sp->outAxis, sp->outSharpness,
                                          sp->outAmplitude);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_sg_specular_forward(const vcl_sg_specular_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->rast && p->normal && p->view && p->roughness && p->outAxis && p->outSharpness && p->outAmplitude);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    vlaunch<KSgSpecFwdF>(k_sg_spec_fwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->rast, p->normal, p->view, p->roughness, p->ior,
                       p->outAxis, p->outSharpness, p->outAmplitude);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_sg_specular_backward(const vcl_sg_specular_params* p, const vcl_sg_specular_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->rast && p->normal && p->view && p->roughness);
    if (!g->roughness) return VCL_OK;
    VCL_REQUIRE(g->outAmplitude && g->outSharpness);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    vlaunch<KSgSpecBwdF>(k_sg_spec_bwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->rast, p->normal, p->view, p->roughness, p->ior,
                       g->outAmplitude, g->outSharpness, g->roughness);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// Workgroups of the lobe backward: one per four 256-pixel chunks (at 1024^2, the 1,024 that are
// resident at once), at most 4,096; with the rotated chunk order of k_sg_bwd_lobes each gets a
// like share of covered pixels, so they end together.  Each adds a row of K x 8 partial sums to
// the column sum.
static unsigned sg_bwd_blocks(size_t P) {
    size_t sweep = (size_t)BS * 4;
    size_t nb = (P + sweep - 1) / sweep;
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

extern "C" size_t vcl_spherical_gaussian_workspace_size(int W, int H, int depth, int channel, int sgnum) {
    (void)channel;
    size_t P = (size_t)W * H * depth;
    size_t cols = (size_t)sgnum * SG_COLS;
    return sizeof(float) * ((size_t)sg_bwd_blocks(P) * cols + colsum_scratch_floats((int)cols)) + 256;
}

extern "C" int vcl_spherical_gaussian_forward(const vcl_spherical_gaussian_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->rast && p->normal && p->diffuse && p->specAxis && p->specSharpness && p->specAmplitude && p->out && p->outDiffenv &&
                p->outSpecenv && p->channel >= 1 && p->channel <= SG_MAXC && p->sgnum >= 0);
    VCL_REQUIRE(p->sgnum == 0 || (p->axis && p->sharpness && p->amplitude));
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    const bool al16 = ((uintptr_t)p->diffuse % 16 == 0) && ((uintptr_t)p->out % 16 == 0) && ((uintptr_t)p->outDiffenv % 16 == 0) &&
                      ((uintptr_t)p->outSpecenv % 16 == 0);
    if (p->sgnum <= SG_MAXK && p->channel == 4 && al16) {
        if (cdiv(P, 4 * BS) >= 4096)
            vlaunch<KSgFwd4F<4 * BS, BS>>(k_sg_fwd4<4 * BS, BS>, dim3(cdiv(P, 4 * BS)), dim3(BS), 0, as_hip(stream), P, p->sgnum, p->rast, p->axis, p->sharpness,
                               p->amplitude, p->normal, p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, p->out, p->outDiffenv,
                               p->outSpecenv);
        else
            vlaunch<KSgFwd4F<2 * BS, 2 * BS>>(k_sg_fwd4<2 * BS, 2 * BS>, dim3(cdiv(P, 2 * BS)), dim3(2 * BS), 0, as_hip(stream), P, p->sgnum, p->rast, p->axis, p->sharpness,
                               p->amplitude, p->normal, p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, p->out, p->outDiffenv,
                               p->outSpecenv);
    }
    else if (p->sgnum <= SG_MAXK)
        vlaunch<KSgFwdFastF>(k_sg_fwd_fast, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->channel, p->sgnum, p->rast, p->axis, p->sharpness,
                           p->amplitude, p->normal, p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, p->out, p->outDiffenv,
                           p->outSpecenv);
    else
        vlaunch<KSgFwdF>(k_sg_fwd, dim3(cdiv(P, BS)), dim3(BS), 0, as_hip(stream), P, p->channel, p->sgnum, p->rast, p->axis, p->sharpness,
                           p->amplitude, p->normal, p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, p->out, p->outDiffenv,
                           p->outSpecenv);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

int vcl_spherical_gaussian_backward_dev(const vcl_spherical_gaussian_params* p, const vcl_spherical_gaussian_grad_params* g,
                                        vcl_sg_lobe_sums* lobes, int gd_fresh, vcl_stream_t stream) {
    if (lobes) lobes->seg = nullptr;
    VCL_REQUIRE(p && g && p->rast && p->normal && p->diffuse && p->specAxis && p->specSharpness && p->specAmplitude && g->out &&
                p->channel >= 1 && p->channel <= SG_MAXC && p->sgnum >= 0);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    int K = p->sgnum, C = p->channel;
    int want = (g->axis || g->sharpness || g->amplitude) && K > 0;
    size_t need = vcl_spherical_gaussian_workspace_size(p->width, p->height, p->depth, C, K);
    if (want && (!g->workspace || g->workspace_size < need)) return VCL_ERR_OOM;
    unsigned nb = sg_bwd_blocks(P);
    size_t cols = (size_t)K * SG_COLS;
    float* part = (float*)g->workspace;
    hipStream_t s = as_hip(stream);
    bool per_pixel = g->normal || g->specAxis || g->specSharpness || g->specAmplitude;
    if (!per_pixel && K <= SG_MAXK && p->outDiffenv) {
        int KP = 1;
        while (KP < K) KP <<= 1;
        vlaunch<KSgBwdLobesF>(k_sg_bwd_lobes, dim3(nb), dim3(BS), 0, s, P, C, K, KP, p->rast, p->axis, p->sharpness, p->amplitude, p->normal,
                           p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, g->out, p->outDiffenv, g->diffuse, part, want, gd_fresh);
    } else {
        size_t lds = sizeof(float) * 4 * (size_t)K * SG_COLS;
        VCL_REQUIRE(lds <= 160 * 1024);
        if (g_vcl_recorder) return VCL_ERR_INVALID;   // per-pixel lobe grads: not a groupable launch
        hipLaunchKernelGGL(k_sg_bwd, dim3(nb), dim3(BS), want ? lds : 0, s, P, C, K, p->rast, p->axis, p->sharpness, p->amplitude, p->normal,
                           p->diffuse, p->specAxis, p->specSharpness, p->specAmplitude, g->out, g->diffuse, g->normal, g->specAxis,
                           g->specSharpness, g->specAmplitude, part, want);
    }
    VCL_LAUNCH_CHECK();
    if (!want) return VCL_OK;
    float* seg = part + (size_t)nb * cols;
    const bool to_caller = lobes && g->axis && g->sharpness && g->amplitude;
    if (lobes) lobes->stage_valid = 0;
    VclColsumStage cs;
    const int nseg = colsum_stage_of(part, (int)nb, (int)cols, seg, cs);
    if (to_caller && lobes->defer_stage) {
        // the caller runs the segment stage (in its texture backward launch) ...
        lobes->stage = cs;
        lobes->stage_valid = 1;
    } else if (colsum_launch(cs, s)) {
        return VCL_ERR_HIP;
    }
    if (to_caller) {
        // ... and its Adam launch adds the sums into the lobe grads
        lobes->seg = seg; lobes->nseg = nseg; lobes->K = K; lobes->C = C;
        return VCL_OK;
    }
    vlaunch<KSgLobeFinalF>(k_sg_lobe_final, dim3(cdiv(cols, 256)), dim3(256), 0, s, seg, nseg, K, C, g->axis, g->sharpness, g->amplitude);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_spherical_gaussian_backward(const vcl_spherical_gaussian_params* p, const vcl_spherical_gaussian_grad_params* g,
                                               vcl_stream_t stream) {
    return vcl_spherical_gaussian_backward_dev(p, g, nullptr, 0, stream);
}

// gout: the backward's output grads (16-byte alignment checked with the input's), or null
static AAArgs aa_args(const vcl_antialias_params* p, const float* gout) {
    AAArgs a;
    a.W = p->width; a.H = p->height; a.depth = p->depth; a.C = p->channel; a.V = p->projNum;
    a.v4 = p->channel == 4 && ((((uintptr_t)p->in) | (uintptr_t)gout) & 15) == 0;
    a.sx = p->s[0]; a.sy = p->s[1];
    a.proj = p->proj; a.idx = p->idx; a.rast = p->rast;
    return a;
}

size_t vcl_antialias_loss_parts(int width, int height, int depth) {
    return (size_t)((width + AA_T - 1) / AA_T) * ((height + AA_T - 1) / AA_T) * (size_t)depth;
}

int vcl_antialias_forward_dev(const vcl_antialias_params* p, const float* target, float* gout, float* part, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->proj && p->idx && p->rast && p->in && p->out && p->channel >= 1 && p->channel <= AA_MAXC && p->depth <= 65535);
    VCL_REQUIRE((!target && !gout && !part) || (target && gout && part));
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    int tilesX = (p->width + AA_T - 1) / AA_T, tilesY = (p->height + AA_T - 1) / AA_T;
    AAArgs a = aa_args(p, nullptr);
    if (gout) a.v4 = a.v4 && ((((uintptr_t)target) | (uintptr_t)gout) & 15) == 0;
    vlaunch<KAaFwdF>(k_aa_fwd, dim3(tilesX * tilesY, p->depth), dim3(BS), 0, as_hip(stream), a, tilesX, p->in, p->out, target, gout, part);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

extern "C" int vcl_antialias_forward(const vcl_antialias_params* p, vcl_stream_t stream) {
    return vcl_antialias_forward_dev(p, nullptr, nullptr, nullptr, stream);
}

extern "C" int vcl_antialias_backward(const vcl_antialias_params* p, const vcl_antialias_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->proj && p->idx && p->rast && p->in && g->out && g->in && p->channel >= 1 && p->channel <= AA_MAXC && p->depth <= 65535);
    size_t P = (size_t)p->width * p->height * p->depth;
    if (!P) return VCL_OK;
    int tilesX = (p->width + AA_T - 1) / AA_T, tilesY = (p->height + AA_T - 1) / AA_T;
    vlaunch<KAaBwdF>(k_aa_bwd, dim3(tilesX * tilesY, p->depth), dim3(BS), 0, as_hip(stream), aa_args(p, g->out), tilesX, p->in, g->out, g->in, g->proj);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// workspace: one partial per block
extern "C" size_t vcl_loss_workspace_size(int size) {
    size_t nb = cdiv((size_t)(size > 0 ? size : 1), (size_t)BS * 4);
    return sizeof(float) * nb + 256;
}

int vcl_mse_loss_grad_dev(const vcl_loss_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->predict && p->target && p->grad && p->buffer && p->loss && p->size >= 0);
    hipStream_t s = as_hip(stream);
    const size_t nb = p->size > 0 ? cdiv((size_t)p->size, (size_t)BS * 4) : 0;
    if (nb) {
        const int vec = ((p->size & 3) == 0 && (((uintptr_t)p->predict | (uintptr_t)p->target | (uintptr_t)p->grad) & 15) == 0) ? 1 : 0;
        vlaunch<KMseF>(k_mse, dim3((unsigned)nb), dim3(BS), 0, s, (size_t)p->size, p->predict, p->target, p->grad, p->buffer, vec);
        VCL_LAUNCH_CHECK();
    }
    return VCL_OK;
}

extern "C" int vcl_mse_loss_backward(const vcl_loss_params* p, vcl_stream_t stream) {
    int st = vcl_mse_loss_grad_dev(p, stream);
    if (st) return st;
    const size_t nb = p->size > 0 ? cdiv((size_t)p->size, (size_t)BS * 4) : 0;
    vlaunch<KMseFinalF>(k_mse_final, dim3(1), dim3(1024), 0, as_hip(stream), p->buffer, nb, p->loss);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}

// ---------------------------------------------------------------- MSE without allocation
// The block partials of vcl_mse_loss_backward_noalloc live in this file-scope device array, one
// copy per GPU, held until the process ends.  Host code never uses the array's name as an
// address (on the host it is only a stand-in for the symbol): each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current.  k_mse writes
// every partial that k_mse_final reads, so nothing needs clearing between calls.  Calls that
// overlap on one GPU share the array and race.
#define VCL_MSE_NOALLOC_PARTIALS (1u << 18)  /* 1 MiB of float partials */
static __device__ __align__(256) float scratch[VCL_MSE_NOALLOC_PARTIALS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(float** out) {
    int dev = 0;
    VCL_HIP_CHECK(hipGetDevice(&dev));
    VCL_REQUIRE(dev >= 0 && dev < kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        VCL_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        VCL_REQUIRE(ptr);
        g_scratch_ptr[dev] = (float*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return VCL_OK;
}
}  // namespace

extern "C" size_t vcl_mse_loss_noalloc_max_size(void) { return (size_t)VCL_MSE_NOALLOC_PARTIALS * BS * 4; }

extern "C" int vcl_mse_loss_backward_noalloc(const vcl_loss_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->predict && p->target && p->grad && p->loss && p->size >= 0);
    if ((size_t)p->size > vcl_mse_loss_noalloc_max_size()) return VCL_ERR_OOM;
    float* part = nullptr;
    int st = scratch_for_current_device(&part);
    if (st) return st;
    vcl_loss_params q = *p;
    q.buffer = part;
    if ((st = vcl_mse_loss_grad_dev(&q, stream))) return st;
    const size_t nb = q.size > 0 ? cdiv((size_t)q.size, (size_t)BS * 4) : 0;
    vlaunch<KMseFinalF>(k_mse_final, dim3(1), dim3(1024), 0, as_hip(stream), part, nb, q.loss);
    VCL_LAUNCH_CHECK();
    return VCL_OK;
}
