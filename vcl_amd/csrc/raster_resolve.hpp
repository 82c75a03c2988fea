// raster_resolve.hpp -- a pixel's winning depth key -> its rasterizer outputs, shared by the
// rasterizer's resolve pass (raster.hip) and the fit step's per-pixel pass that does the
// resolve itself (ops.hip).  Every function evaluates its float expressions without FP
// contraction (the CPU oracle's operation order), so both give identical results.
#pragma once
#include "common.hpp"

// Depth planes of the original triangle (see the oracle's depth_planes: same float ops):
// Az Bz Cz Aw Bw Cw Ad Bd Cd.
__device__ __forceinline__ void depth_planes(float4 c0, float4 c1, float4 c2, float* zp) {
#pragma clang fp contract(off)
    float4 c[3] = {c0, c1, c2};
    float a[3], b[3], d[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float4 j = c[(i + 1) % 3], k = c[(i + 2) % 3];
        a[i] = j.y * k.w - j.w * k.y;
        b[i] = j.w * k.x - j.x * k.w;
        d[i] = j.x * k.y - j.y * k.x;
    }
    zp[0] = (c0.z * a[0] + c1.z * a[1]) + c2.z * a[2];
    zp[1] = (c0.z * b[0] + c1.z * b[1]) + c2.z * b[2];
    zp[2] = (c0.z * d[0] + c1.z * d[1]) + c2.z * d[2];
    zp[3] = (c0.w * a[0] + c1.w * a[1]) + c2.w * a[2];
    zp[4] = (c0.w * b[0] + c1.w * b[1]) + c2.w * b[2];
    zp[5] = (c0.w * d[0] + c1.w * d[1]) + c2.w * d[2];
    zp[6] = (a[0] + a[1]) + a[2];
    zp[7] = (b[0] + b[1]) + b[2];
    zp[8] = (d[0] + d[1]) + d[2];
}

// Pixel-centre NDC coordinate of column (row) i for the step xs = 2 / W (ys = 2 / H).
__device__ __forceinline__ float pixel_centre(int i, float xs) {
#pragma clang fp contract(off)
    return (float)((i + 0.5) * (double)xs - 1.0);
}

// The winning key's triangle t (its clip-space vertices c0, c1, c2) at the pixel centre
// (pixx, pixy): rast = (u, v, z/w, t + 1) by the perspective closed form of rasterize.cu:101-121,
// and, when want_db, the analytic screen derivatives db = (du/dx, du/dy, dv/dx, dv/dy).
__device__ __forceinline__ void resolve_pixel(float4 c0, float4 c1, float4 c2, u32 t, float xs, float ys, float pixx, float pixy, bool want_db,
                                              float4& o, float4& d) {
#pragma clang fp contract(off)
    float w0 = c0.w, w1 = c1.w, w2 = c2.w;
    float P0x = w0 * pixx - c0.x, P0y = w0 * pixy - c0.y;
    float P1x = w1 * pixx - c1.x, P1y = w1 * pixy - c1.y;
    float P2x = w2 * pixx - c2.x, P2y = w2 * pixy - c2.y;
    float E0x = P0x - P2x, E0y = P0y - P2y, E1x = P1x - P2x, E1y = P1y - P2y;
    float D = E0x * E1y - E0y * E1x;
    float u = (P1x * P2y - P1y * P2x) / D;
    float v = (P2x * P0y - P2y * P0x) / D;
    float zp[9];
    depth_planes(c0, c1, c2, zp);
    float zz = (zp[0] * pixx + zp[1] * pixy + zp[2]) / (zp[3] * pixx + zp[4] * pixy + zp[5]);
    o = make_float4(u, v, zz, (float)(t + 1));
    d = make_float4(0.f, 0.f, 0.f, 0.f);
    if (want_db) {
        float dNuX = xs * (w1 * P2y - w2 * P1y), dNuY = ys * (w2 * P1x - w1 * P2x);
        float dNvX = xs * (w2 * P0y - w0 * P2y), dNvY = ys * (w0 * P2x - w2 * P0x);
        float dDX = xs * ((w0 - w2) * E1y - (w1 - w2) * E0y);
        float dDY = ys * ((w1 - w2) * E0x - (w0 - w2) * E1x);
        d = make_float4((dNuX - u * dDX) / D, (dNuY - u * dDY) / D, (dNvX - v * dDX) / D, (dNvY - v * dDY) / D);
    }
}

// No winner (the key buffer's all-ones start value).
constexpr unsigned long long RAST_NO_KEY = ~0ull;
