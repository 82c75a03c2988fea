// extra.hip -- the reference ops beyond the fit chain, for gfx950: noise, Gaussian filter, SSIM
// and texture losses, SGD / Nadam and the optimizer helpers, buffer utilities, texture bilinear
// downsampling, SG bake / normalize / disperse, ComputeNormal (forward and a real backward), and
// the debug raster shadings.  Each op is HBM- or latency-bound streaming work: one thread per
// element or pixel, coalesced along rows, window taps staged in LDS where a pixel reads many
// neighbours (SSIM), reductions in a fixed order (no float atomics) except the vertex scatter
// of ComputeNormal.
#include "internal.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int BS = 256;

inline unsigned grid_for(size_t n) {
    size_t b = (n + BS - 1) / BS;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 1u << 20));
}

// getUniform's integer mix (common.h:194-204): Bob Jenkins' 96-bit mix() of lookup2.c.
__device__ __forceinline__ void jenkins_mix(u32& a, u32& b, u32& c) {
    a -= b; a -= c; a ^= (c >> 13);
    b -= c; b -= a; b ^= (a << 8);
    c -= a; c -= b; c ^= (b >> 13);
    a -= b; a -= c; a ^= (c >> 12);
    b -= c; b -= a; b ^= (a << 16);
    c -= a; c -= b; c ^= (b >> 5);
    a -= b; a -= c; a ^= (c >> 3);
    b -= c; b -= a; b ^= (a << 10);
    c -= a; c -= b; c ^= (b >> 15);
}
__device__ __forceinline__ float uniform01(u32 a, u32 b, u32 c) {
    jenkins_mix(a, b, c);
    return (float)(c & 0x007fffffu) / (float)0x007fffff;
}
__device__ __forceinline__ float clampf(float a, float lo, float hi) { return a < lo ? lo : a > hi ? hi : a; }

// Block sum of one float per thread (fixed order: wave sums, then the waves in order).
__device__ __forceinline__ float block_sum(float v, float* sh) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.f;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[w];
    __syncthreads();
    return s;
}

// One workgroup sums n float partials in a fixed order into *out (scaled by `scale`).
__global__ void __launch_bounds__(1024) k_sum_partials(const float* part, size_t n, double scale, float* out) {
    __shared__ double sh[1024];
    double s = 0.0;
    size_t per = (n + 1023) / 1024, b = threadIdx.x * per, e = std::min(n, b + per);
    for (size_t i = b; i < e; i++) s += (double)part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(sh[0] * scale);
}

// ---------------------------------------------------------------- noise
__global__ void __launch_bounds__(BS) k_noise(size_t P, int C, float intensity, const float* in, float* out, u32 seed) {
#pragma clang fp contract(off)
    for (size_t p = (size_t)blockIdx.x * BS + threadIdx.x; p < P; p += (size_t)gridDim.x * BS)
        for (int i = 0; i < C; i++) {
            float x = in[p * C + i];
            out[p * C + i] = x + intensity * (uniform01((u32)p, seed + (u32)i, 0xcafef00du) - x);
        }
}

// ---------------------------------------------------------------- optimizers
__global__ void __launch_bounds__(BS) k_sgd(size_t n, float* param, const float* grad, double eta) {
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS)
        add_nan_check(param[i], (float)(-eta * (double)grad[i]));
}

struct NadamArgs {
    double mu, mu0, mu1, rho, alpha, mut0, mut1, rhot, eps;
};
__global__ void __launch_bounds__(BS) k_nadam(size_t n, float* param, const float* grad, float* m, float* v, NadamArgs a) {
#pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
        double g = grad[i];
        float mi = (float)(a.mu * (double)m[i] + (1 - a.mu) * g);
        float vi = (float)(a.rho * (double)v[i] + (1 - a.rho) * g * g);
        m[i] = mi;
        v[i] = vi;
        double mh = a.mu1 / (1 - a.mut1) * (double)mi + (1 - a.mu0) / (1 - a.mut0) * g;
        double vh = 1 / (1 - a.rhot) * (double)vi;
        add_nan_check(param[i], (float)(-mh * a.alpha / (sqrt(vh) + a.eps)));
    }
}

__global__ void __launch_bounds__(BS) k_random_fill(size_t n, float* dst, u32 seed, u32 salt, float lo, float hi, int add) {
#pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
        float r = lo + (hi - lo) * uniform01((u32)i, seed, salt);
        dst[i] = add ? dst[i] + r : r;
    }
}

// ---------------------------------------------------------------- buffer utilities
__global__ void __launch_bounds__(BS) k_linear(size_t n, float* dst, const float* src, float w, float b) {
#pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) dst[i] = src[i] * w + b;
}
__global__ void __launch_bounds__(BS) k_zero(size_t n, float* dst) {
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) dst[i] = 0.f;
}
__global__ void __launch_bounds__(BS) k_clamp_buf(size_t n, float* dst, const float* src, float lo, float hi) {
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) dst[i] = clampf(src[i], lo, hi);
}
__global__ void __launch_bounds__(BS) k_step_buf(size_t n, float* dst, const float* src, float thr) {
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) dst[i] = src[i] < thr ? 0.f : 1.f;
}
__global__ void __launch_bounds__(BS) k_normalize_rows(int rows, int dim, float* dst, const float* src) {
#pragma clang fp contract(off)
    int r = blockIdx.x * BS + threadIdx.x;
    if (r >= rows) return;
    float l2 = 0.f;
    for (int i = 0; i < dim; i++) { float x = src[(size_t)r * dim + i]; l2 += x * x; }
    if (l2 <= 0.f) return;
    float il = 1.f / sqrtf(l2);
    for (int i = 0; i < dim; i++) dst[(size_t)r * dim + i] *= il;
}
struct Value16 { float v[16]; };
__global__ void __launch_bounds__(BS) k_set_value(int rows, int dim, float* dst, Value16 val) {
    int r = blockIdx.x * BS + threadIdx.x;
    if (r >= rows) return;
    for (int i = 0; i < dim; i++) dst[(size_t)r * dim + i] = val.v[i];
}
__global__ void __launch_bounds__(BS) k_spherical_random(int n, float* dst, u32 seed) {
#pragma clang fp contract(off)
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    float theta = (uniform01((u32)i, seed, 0xba5eba11u) * 2.f - 1.f) * 3.14159265f;
    float z = uniform01((u32)i, seed, 0xdeadba11u) * 2.f - 1.f;
    float c = sqrtf(1.f - z * z);
    dst[3 * i] = c * cosf(theta);
    dst[3 * i + 1] = c * sinf(theta);
    dst[3 * i + 2] = z;
}
// mean row distance in one workgroup: each thread sums a contiguous share in order
__global__ void __launch_bounds__(1024) k_distance_error(int num, int dim, const float* a, const float* b, float* out) {
#pragma clang fp contract(off)
    __shared__ double sh[1024];
    int per = (num + 1023) / 1024, r0 = threadIdx.x * per, r1 = min(num, r0 + per);
    double s = 0.0;
    for (int r = r0; r < r1; r++) {
        float d = 0.f;
        for (int k = 0; k < dim; k++) { float e = a[(size_t)r * dim + k] - b[(size_t)r * dim + k]; d += e * e; }
        s += (double)sqrtf(d);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(sh[0] / num);
}

// Texture::bilinearDownsampling's level step (buffer.cu:410-435): dst (w x h) from src (2w wide)
__global__ void __launch_bounds__(BS) k_bilinear_down(const float* src, float* dst, int w, int h, int C) {
#pragma clang fp contract(off)
    int p = blockIdx.x * BS + threadIdx.x;
    if (p >= w * h) return;
    const float filt[4] = {.125f, .375f, .375f, .125f};
    int px = p % w, py = p / w;
    int x2 = px << 1, y2 = py << 1, W2 = w << 1, H2 = h << 1;
    int xs = x2 > 0 ? -1 : 0, xe = x2 < W2 - 2 ? 2 : 1;
    int ys = y2 > 0 ? -1 : 0, ye = y2 < H2 - 2 ? 2 : 1;
    int nc = C < 3 ? C : 3;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int x = xs; x <= xe; x++)
        for (int y = ys; y <= ye; y++) {
            float f = filt[x + 1] * filt[y + 1];
            size_t q = (size_t)(x2 + x) + (size_t)W2 * (y2 + y);
            for (int i = 0; i < nc; i++) acc[i] += src[q * C + i] * f;
        }
    for (int i = 0; i < nc; i++) dst[(size_t)p * C + i] = acc[i];
}

// SGBuffer::bake into level 0: one thread per texel, lobes staged through LDS in chunks
constexpr int BAKE_CHUNK = 256;
__global__ void __launch_bounds__(BS) k_sg_bake(int num, int C, const float* axis, const float* sharp, const float* amp, int W, int H,
                                                float* tex0) {
#pragma clang fp contract(off)
    __shared__ float s_ax[BAKE_CHUNK * 3], s_sh[BAKE_CHUNK];
    int p = blockIdx.x * BS + threadIdx.x;
    bool live = p < W * H;
    int px = live ? p % W : 0, py = live ? p / W : 0;
    float phi = (1.f - 2.f * (float)px / (float)W) * 3.14159265f;
    float theta = (2.f * (float)py / (float)H - 1.f) * 1.5707963f;
    float ax = sinf(phi) * cosf(theta), ay = sinf(theta), az = cosf(phi) * cosf(theta);
    for (int c0 = 0; c0 < num; c0 += BAKE_CHUNK) {
        int n = min(BAKE_CHUNK, num - c0);
        for (int i = threadIdx.x; i < n; i += BS) {
            s_ax[3 * i] = axis[3 * (size_t)(c0 + i)];
            s_ax[3 * i + 1] = axis[3 * (size_t)(c0 + i) + 1];
            s_ax[3 * i + 2] = axis[3 * (size_t)(c0 + i) + 2];
            s_sh[i] = sharp[c0 + i];
        }
        __syncthreads();
        if (live)
            for (int i = 0; i < n; i++) {
                float d = ax * s_ax[3 * i] + ay * s_ax[3 * i + 1] + az * s_ax[3 * i + 2];
                float v = expf(s_sh[i] * (d - 1.f));
                for (int k = 0; k < C; k++) add_nan_check(tex0[(size_t)p * C + k], amp[(size_t)(c0 + i) * C + k] * v);
            }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(BS) k_sg_normalize(int num, int C, float* axis, float* sharp, float* amp) {
#pragma clang fp contract(off)
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= num) return;
    float a0 = axis[3 * i], a1 = axis[3 * i + 1], a2 = axis[3 * i + 2];
    float l2 = a0 * a0 + a1 * a1 + a2 * a2;
    if (l2 > 0.f) {
        float il = 1.f / sqrtf(l2);
        axis[3 * i] = a0 * il; axis[3 * i + 1] = a1 * il; axis[3 * i + 2] = a2 * il;
    }
    sharp[i] = clampf(sharp[i], 0.f, 8.f);
    for (int k = 0; k < C; k++) amp[(size_t)i * C + k] = clampf(amp[(size_t)i * C + k], 0.f, 1.f);
}

__global__ void __launch_bounds__(BS) k_sg_disperse(int num, int C, const float* axis, const float* sharp, const float* amp, float* buffer) {
#pragma clang fp contract(off)
    int p = blockIdx.x * BS + threadIdx.x;
    if (p >= num) return;
    float a0 = axis[3 * p], a1 = axis[3 * p + 1], a2 = axis[3 * p + 2], s = sharp[p];
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    for (int i = 0; i < num; i++) {
        float sq = 0.f;
        for (int k = 0; k < C; k++) { float d = amp[(size_t)p * C + k] - amp[(size_t)i * C + k]; sq += d * d; }
        float v0 = a0 - axis[3 * i], v1 = a1 - axis[3 * i + 1], v2 = a2 - axis[3 * i + 2];
        float l2 = (float)fmax((double)(v0 * v0 + v1 * v1 + v2 * v2), 1e-6);
        float il = 1.f / sqrtf(l2);
        v0 *= il; v1 *= il; v2 *= il;
        float w = (float)(1.0 / fmax((double)(l2 * s * sq), 1e-6));
        f0 += v0 * w; f1 += v1 * w; f2 += v2 * w;
    }
    buffer[3 * p] = f0; buffer[3 * p + 1] = f1; buffer[3 * p + 2] = f2;
}
__global__ void __launch_bounds__(BS) k_add3(int num, float* axis, const float* buffer) {
    int p = blockIdx.x * BS + threadIdx.x;
    if (p >= 3 * num) return;
    axis[p] += buffer[p];
}

// ---------------------------------------------------------------- texture loss
__global__ void __launch_bounds__(BS) k_texture_loss(int W, int H, int D, const float* pred, const float* target, float* buffer, float* part) {
#pragma clang fp contract(off)
    __shared__ float sh[BS / 64];
    size_t n = (size_t)W * H * D;
    size_t p = (size_t)blockIdx.x * BS + threadIdx.x;
    float v = 0.f;
    if (p < n) {
        int pz = (int)(p % D), pix = (int)(p / D), px = pix % W, py = pix / W;
        int w = W / 4, w2 = W / 2;
        if (!((px < w || w2 <= px) && (py < w || w2 <= py))) { float t = pred[p] - target[p]; v = t * t; }
        buffer[p] = v;
    }
    float s = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------- SSIM
// A 16 x 16 tile of pixels, one thread each, with its 26 x 26 neighbourhood of one value per
// pixel (one channel at a time) staged in LDS for both images.  Window sums run over ix (outer)
// then iy (inner) as the reference's do.
constexpr int SS_T = 16, SS_K = 5, SS_E = SS_T + 2 * SS_K;

struct SsimArgs {
    int W, H, D, mode;
    const float* pred;
    const float* target;
    float* grad;
    float* map;      // 1 - S per element
    float* alpha;    // exact mode: per-element window coefficients
    float* beta;
    float* gam;
    float* part;     // per-workgroup sums of the map
};

__device__ __forceinline__ void ssim_stage(const SsimArgs& a, const float* src, float* lds, int x0, int y0, int c) {
    for (int i = threadIdx.x; i < SS_E * SS_E; i += BS) {
        int lx = i % SS_E, ly = i / SS_E, gx = x0 - SS_K + lx, gy = y0 - SS_K + ly;
        lds[i] = (gx >= 0 && gx < a.W && gy >= 0 && gy < a.H) ? src[((size_t)gx + (size_t)a.W * gy) * a.D + c] : 0.f;
    }
}

__global__ void __launch_bounds__(BS) k_ssim_stats(SsimArgs a) {
#pragma clang fp contract(off)
    __shared__ float sx[SS_E * SS_E], sy[SS_E * SS_E];
    __shared__ float sh[BS / 64];
    const int tilesX = (a.W + SS_T - 1) / SS_T;
    const int x0 = (blockIdx.x % tilesX) * SS_T, y0 = (blockIdx.x / tilesX) * SS_T;
    const int tx = threadIdx.x % SS_T, ty = threadIdx.x / SS_T;
    const int px = x0 + tx, py = y0 + ty;
    const bool live = px < a.W && py < a.H;
    const float c1 = 1e-4f, c2 = 1e-3f, inv = 1.f / (float)(SS_K * SS_K);
    const int sx0 = max(px - SS_K, 0), ex0 = min(px + SS_K, a.W - 1);
    const int sy0 = max(py - SS_K, 0), ey0 = min(py + SS_K, a.H - 1);
    float total = 0.f;
    for (int c = 0; c < a.D; c++) {
        ssim_stage(a, a.pred, sx, x0, y0, c);
        ssim_stage(a, a.target, sy, x0, y0, c);
        __syncthreads();
        if (live) {
            float ax = 0.f, ay = 0.f;
            for (int ix = sx0; ix <= ex0; ix++)
                for (int iy = sy0; iy <= ey0; iy++) {
                    int l = (ix - x0 + SS_K) + SS_E * (iy - y0 + SS_K);
                    ax += sx[l]; ay += sy[l];
                }
            float mx = ax * inv, my = ay * inv;
            float vx = 0.f, vy = 0.f, cv = 0.f;
            for (int ix = sx0; ix <= ex0; ix++)
                for (int iy = sy0; iy <= ey0; iy++) {
                    int l = (ix - x0 + SS_K) + SS_E * (iy - y0 + SS_K);
                    float dx = sx[l] - mx, dy = sy[l] - my;
                    vx += dx * dx; vy += dy * dy; cv += dx * dy;
                }
            vx *= inv; vy *= inv; cv *= inv;
            float n1 = 2.f * mx * my + c1, d1 = mx * mx + my * my + c1;
            float n2 = 2.f * cv + c2, d2 = vx + vy + c2;
            float dinv = 1.f / (d1 * d2);
            float S = n1 * n2 * dinv;
            size_t p = ((size_t)px + (size_t)a.W * py) * a.D + c;
            float m = 1.f - S;
            a.map[p] = m;
            total += m;
            int lc = (tx + SS_K) + SS_E * (ty + SS_K);
            float xp = sx[lc], yp = sy[lc];
            if (a.mode == VCL_SSIM_GRAD_REF) {
                a.grad[p] = -(my * n2 + n1 * (yp - my) - m * (mx * d2 + d1 * (xp - mx))) * 2.f * inv * dinv;
            } else {
                int n = (ex0 - sx0 + 1) * (ey0 - sy0 + 1);
                float A = 2.f * (my * n2 * dinv - S * mx / d1);
                float B = -S / d2;
                float Cc = 2.f * n1 * dinv;
                float exx = inv * (ax - (float)n * mx), eyy = inv * (ay - (float)n * my);
                a.alpha[p] = A - 2.f * B * (mx + exx) - Cc * (my + eyy);
                a.beta[p] = 2.f * B;
                a.gam[p] = Cc;
            }
        }
        __syncthreads();
    }
    float s = block_sum(total, sh);
    if (threadIdx.x == 0) a.part[blockIdx.x] = s;
}

// exact-mode gradient: box sums of the three coefficient maps over each element's window
__global__ void __launch_bounds__(BS) k_ssim_grad(SsimArgs a) {
#pragma clang fp contract(off)
    __shared__ float s_a[SS_E * SS_E], s_b[SS_E * SS_E], s_g[SS_E * SS_E];
    const int tilesX = (a.W + SS_T - 1) / SS_T;
    const int x0 = (blockIdx.x % tilesX) * SS_T, y0 = (blockIdx.x / tilesX) * SS_T;
    const int tx = threadIdx.x % SS_T, ty = threadIdx.x / SS_T;
    const int px = x0 + tx, py = y0 + ty;
    const bool live = px < a.W && py < a.H;
    const float inv = 1.f / (float)(SS_K * SS_K);
    const int sx0 = max(px - SS_K, 0), ex0 = min(px + SS_K, a.W - 1);
    const int sy0 = max(py - SS_K, 0), ey0 = min(py + SS_K, a.H - 1);
    for (int c = 0; c < a.D; c++) {
        ssim_stage(a, a.alpha, s_a, x0, y0, c);
        ssim_stage(a, a.beta, s_b, x0, y0, c);
        ssim_stage(a, a.gam, s_g, x0, y0, c);
        __syncthreads();
        if (live) {
            float ta = 0.f, tb = 0.f, tg = 0.f;
            for (int ix = sx0; ix <= ex0; ix++)
                for (int iy = sy0; iy <= ey0; iy++) {
                    int l = (ix - x0 + SS_K) + SS_E * (iy - y0 + SS_K);
                    ta += s_a[l]; tb += s_b[l]; tg += s_g[l];
                }
            size_t q = ((size_t)px + (size_t)a.W * py) * a.D + c;
            a.grad[q] = -inv * (ta + tb * a.pred[q] + tg * a.target[q]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- filter
// Scratch layout of the 256-float filter buffers: taps 0..2k (2k + 1 <= 255), slot 255 = k for
// the forward weights, and the second moment m2 for the backward's i^2-weighted taps.
constexpr int FILTER_SLOTS = 256;
constexpr int FILTER_MAXK = 127;

// Filter::forward's host weights (filter.cu:66-79) on the device, in the same order.
__global__ void k_filter_weights(const float* sigma, float* w, float* g) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s_ = *sigma;
    int k = (int)ceilf(s_ * 4.f);
    if (k * 2 > 255) k = 255 / 2;
    if (k < 0) k = 0;
    float s2 = .5f / (s_ * s_), s = 0.f;
    for (int i = -k; i <= k; i++) { w[i + k] = expf((float)(-i * i) * s2); s += w[i + k]; }
    s = 1.f / s;
    float m2 = 0.f;
    for (int i = -k; i <= k; i++) { w[i + k] *= s; m2 += w[i + k] * (float)(i * i); }
    w[FILTER_SLOTS - 1] = (float)k;
    if (g) {
        for (int i = -k; i <= k; i++) g[i + k] = w[i + k] * (float)(i * i);
        g[FILTER_SLOTS - 1] = m2;
    }
}

struct FilterArgs {
    int W, H, D, C;
};

// one pass along x (AXIS 0) or y (AXIS 1); up to two weight sets on the same input
template <int AXIS, int NW>
__global__ void __launch_bounds__(BS) k_filter_pass(FilterArgs f, const float* wts0, const float* wts1, const float* in, float* out0,
                                                    float* out1) {
#pragma clang fp contract(off)
    __shared__ float w0[FILTER_SLOTS], w1[FILTER_SLOTS];
    for (int i = threadIdx.x; i < FILTER_SLOTS; i += BS) { w0[i] = wts0[i]; if (NW > 1) w1[i] = wts1[i]; }
    __syncthreads();
    const int k = (int)wts0[FILTER_SLOTS - 1];
    const size_t P = (size_t)f.W * f.H * f.D;
    for (size_t p = (size_t)blockIdx.x * BS + threadIdx.x; p < P; p += (size_t)gridDim.x * BS) {
        const int px = (int)(p % f.W), py = (int)((p / f.W) % f.H);
        const size_t layer = p / ((size_t)f.W * f.H);
        const int pos = AXIS == 0 ? px : py, len = AXIS == 0 ? f.W : f.H;
        int s = pos - k, f0 = 0;
        if (s < 0) { f0 = -s; s = 0; }
        const int e = min(pos + k, len - 1);
        for (int c = 0; c < f.C; c++) {
            float a0 = 0.f, a1 = 0.f;
            for (int i = s, fi = f0; i <= e; i++, fi++) {
                size_t q = AXIS == 0 ? (size_t)i + (size_t)f.W * (py + (size_t)f.H * layer) : (size_t)px + (size_t)f.W * (i + (size_t)f.H * layer);
                float x = in[q * f.C + c];
                a0 += x * w0[fi];
                if (NW > 1) a1 += x * w1[fi];
            }
            out0[p * f.C + c] = a0;
            if (NW > 1) out1[p * f.C + c] = a1;
        }
    }
}

// the backward's vertical passes: gin = V_w(hf) (stored), and the sigma term per element,
// in * ((V_g(hf) + V_w(hg)) - 2 m2 gin), summed per workgroup into part
__global__ void __launch_bounds__(BS) k_filter_bwd_v(FilterArgs f, const float* wts, const float* gwts, const float* hf, const float* hg,
                                                     const float* in, float* gin, float* part) {
#pragma clang fp contract(off)
    __shared__ float w[FILTER_SLOTS], g[FILTER_SLOTS];
    __shared__ float sh[BS / 64];
    for (int i = threadIdx.x; i < FILTER_SLOTS; i += BS) { w[i] = wts[i]; g[i] = gwts[i]; }
    __syncthreads();
    const int k = (int)w[FILTER_SLOTS - 1];
    const float m2 = g[FILTER_SLOTS - 1];
    const size_t P = (size_t)f.W * f.H * f.D;
    float term = 0.f;
    const size_t p = (size_t)blockIdx.x * BS + threadIdx.x;
    if (p < P) {
        const int px = (int)(p % f.W), py = (int)((p / f.W) % f.H);
        const size_t layer = p / ((size_t)f.W * f.H);
        int s = py - k, f0 = 0;
        if (s < 0) { f0 = -s; s = 0; }
        const int e = min(py + k, f.H - 1);
        for (int c = 0; c < f.C; c++) {
            float gi = 0.f, vg = 0.f, vfg = 0.f;
            for (int i = s, fi = f0; i <= e; i++, fi++) {
                size_t q = ((size_t)px + (size_t)f.W * (i + (size_t)f.H * layer)) * f.C + c;
                float a = hf[q], b = hg[q];
                gi += a * w[fi];
                vg += a * g[fi];
                vfg += b * w[fi];
            }
            gin[p * f.C + c] = gi;
            term += in[p * f.C + c] * ((vg + vfg) - 2.f * m2 * gi);
        }
    }
    float sum = block_sum(term, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// d loss / d sigma = (sum of the partials) / sigma^3
__global__ void __launch_bounds__(1024) k_filter_sigma(const float* part, size_t n, const float* sigma, float* gsigma) {
    __shared__ double sh[1024];
    double s = 0.0;
    size_t per = (n + 1023) / 1024, b = threadIdx.x * per, e = std::min(n, b + per);
    for (size_t i = b; i < e; i++) s += (double)part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sg = (double)*sigma;
        *gsigma = (float)(sh[0] / (sg * sg * sg));
    }
}

// ---------------------------------------------------------------- ComputeNormal
__device__ __forceinline__ void cn_corner(const float* a, const float* b, const float* c, int mode, float& th, float* dA, float* dB, float* dC) {
#pragma clang fp contract(off)
    float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float d = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    if (mode == VCL_NORMAL_ANGLE_REF) {
        th = acosf(d);
        if (dA) {
            float s = -1.f / sqrtf(1.f - d * d);
            for (int i = 0; i < 3; i++) { dB[i] = s * e2[i]; dC[i] = s * e1[i]; dA[i] = -(dB[i] + dC[i]); }
        }
    } else {
        float l1 = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]), l2 = sqrtf(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        float cs = d / (l1 * l2);
        th = acosf(cs);
        if (dA) {
            float s = -1.f / sqrtf(1.f - cs * cs);
            for (int i = 0; i < 3; i++) {
                float db = e2[i] / (l1 * l2) - cs * e1[i] / (l1 * l1);
                float dc = e1[i] / (l1 * l2) - cs * e2[i] / (l2 * l2);
                dB[i] = s * db; dC[i] = s * dc; dA[i] = -(dB[i] + dC[i]);
            }
        }
    }
}
__device__ __forceinline__ float cn_face(const float* v0, const float* v1, const float* v2, float* n) {
#pragma clang fp contract(off)
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]}, e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float len = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    float il = 1.f / len;
    n[0] = c[0] * il; n[1] = c[1] * il; n[2] = c[2] * il;
    return len;
}
__device__ __forceinline__ void load3(const float* p, u32 i, float* o) { o[0] = p[3 * (size_t)i]; o[1] = p[3 * (size_t)i + 1]; o[2] = p[3 * (size_t)i + 2]; }

__global__ void __launch_bounds__(BS) k_cn_scatter(int Tn, const float* pos, const u32* idx, int mode, float* sum) {
#pragma clang fp contract(off)
    int t = blockIdx.x * BS + threadIdx.x;
    if (t >= Tn) return;
    u32 f[3] = {idx[3 * (size_t)t], idx[3 * (size_t)t + 1], idx[3 * (size_t)t + 2]};
    float p[3][3];
    for (int c = 0; c < 3; c++) load3(pos, f[c], p[c]);
    float n[3];
    cn_face(p[0], p[1], p[2], n);
    for (int c = 0; c < 3; c++) {
        float th;
        cn_corner(p[c], p[(c + 1) % 3], p[(c + 2) % 3], mode, th, nullptr, nullptr, nullptr);
        for (int i = 0; i < 3; i++) atomicAdd(&sum[3 * (size_t)f[c] + i], n[i] * th);
    }
}
__global__ void __launch_bounds__(BS) k_cn_normalize(int V, const float* sum, float* out) {
#pragma clang fp contract(off)
    int v = blockIdx.x * BS + threadIdx.x;
    if (v >= V) return;
    float s[3];
    load3(sum, (u32)v, s);
    float il = 1.f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int i = 0; i < 3; i++) out[3 * (size_t)v + i] = s[i] * il;
}
// backward per triangle: the normalisation's grad of each corner's vertex, then through the
// face normal and the corner angles into the three positions
__global__ void __launch_bounds__(BS) k_cn_backward(int Tn, const float* pos, const u32* idx, int mode, const float* sum, const float* out,
                                                    const float* gout, float* gpos) {
#pragma clang fp contract(off)
    int t = blockIdx.x * BS + threadIdx.x;
    if (t >= Tn) return;
    u32 f[3] = {idx[3 * (size_t)t], idx[3 * (size_t)t + 1], idx[3 * (size_t)t + 2]};
    float p[3][3];
    for (int c = 0; c < 3; c++) load3(pos, f[c], p[c]);
    float n[3];
    float len = cn_face(p[0], p[1], p[2], n);
    float a[3] = {0.f, 0.f, 0.f}, gp[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    for (int c = 0; c < 3; c++) {
        float s[3], o[3], g[3];
        load3(sum, f[c], s); load3(out, f[c], o); load3(gout, f[c], g);
        float l = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        float go = g[0] * o[0] + g[1] * o[1] + g[2] * o[2];
        float G[3] = {(g[0] - go * o[0]) / l, (g[1] - go * o[1]) / l, (g[2] - go * o[2]) / l};
        float th, dA[3], dB[3], dC[3];
        cn_corner(p[c], p[(c + 1) % 3], p[(c + 2) % 3], mode, th, dA, dB, dC);
        for (int i = 0; i < 3; i++) a[i] += th * G[i];
        float b = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
        for (int i = 0; i < 3; i++) { gp[c][i] += b * dA[i]; gp[(c + 1) % 3][i] += b * dB[i]; gp[(c + 2) % 3][i] += b * dC[i]; }
    }
    float an = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
    float h[3] = {(a[0] - an * n[0]) / len, (a[1] - an * n[1]) / len, (a[2] - an * n[2]) / len};
    for (int c = 0; c < 3; c++) {
        const float* pb = p[(c + 1) % 3];
        const float* pc = p[(c + 2) % 3];
        float d[3] = {pb[0] - pc[0], pb[1] - pc[1], pb[2] - pc[2]};
        gp[c][0] += d[1] * h[2] - d[2] * h[1];
        gp[c][1] += d[2] * h[0] - d[0] * h[2];
        gp[c][2] += d[0] * h[1] - d[1] * h[0];
    }
    for (int c = 0; c < 3; c++)
        for (int i = 0; i < 3; i++) atomicAdd(&gpos[3 * (size_t)f[c] + i], gp[c][i]);
}

// ---------------------------------------------------------------- debug raster shadings
__global__ void __launch_bounds__(BS) k_idhash(size_t P, const float* rast, float* out) {
    size_t p = (size_t)blockIdx.x * BS + threadIdx.x;
    if (p >= P) return;
    int id = (int)rast[4 * p + 3] - 1;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (id >= 0) {
        u32 a = (u32)id, b = 0xbeefcadeu, c = 0xdeadcafeu;
        jenkins_mix(a, b, c);
        const float d = (float)0x007fffff;
        o = make_float4((float)(a & 0x007fffffu) / d, (float)(b & 0x007fffffu) / d, (float)(c & 0x007fffffu) / d, 1.f);
    }
    ((float4*)out)[p] = o;
}
__global__ void __launch_bounds__(BS) k_wireframe(size_t P, const float* rast, const float* db, float* out) {
#pragma clang fp contract(off)
    size_t p = (size_t)blockIdx.x * BS + threadIdx.x;
    if (p >= P) return;
    float4 r = ((const float4*)rast)[p];
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((int)r.w - 1 >= 0) {
        float4 g = ((const float4*)db)[p];
        float var[3] = {r.x, r.y, 1.f - r.x - r.y};
        float d[3] = {0.5f * (fabsf(g.x) + fabsf(g.y)), 0.5f * (fabsf(g.z) + fabsf(g.w)), 0.5f * (fabsf(g.x + g.z) + fabsf(g.y + g.w))};
        float fm = 1e30f;
        for (int i = 0; i < 3; i++) fm = fminf(fm, 1.f / (1.f + expf(16.f * (d[i] - var[i]))));
        float e = 1.f - fm;
        o = make_float4(e, e, e, 1.f);
    }
    ((float4*)out)[p] = o;
}

}  // namespace

// =============================================================================== C ABI
#define LAUNCH(kernel, grid, block, stream, ...)                                     \
    do {                                                                             \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, __VA_ARGS__); \
        VCL_LAUNCH_CHECK();                                                          \
    } while (0)

extern "C" int vcl_noise_forward(const vcl_noise_params* p, uint32_t seed, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->in && p->out && p->width > 0 && p->height > 0 && p->depth > 0 && p->channel > 0);
    size_t P = (size_t)p->width * p->height * p->depth;
    LAUNCH(k_noise, grid_for(P), BS, as_hip(stream), P, p->channel, p->intensity, p->in, p->out, (u32)seed);
    return VCL_OK;
}

extern "C" int vcl_sgd_step(const vcl_optimize_params* p, const vcl_sgd_hyper_params* h, vcl_stream_t stream) {
    VCL_REQUIRE(p && h && p->param && p->grad && p->size >= 0);
    if (p->size == 0) return VCL_OK;
    LAUNCH(k_sgd, grid_for((size_t)p->size), BS, as_hip(stream), (size_t)p->size, p->param, p->grad, h->eta);
    return VCL_OK;
}

extern "C" int vcl_nadam_step(const vcl_optimize_params* p, const vcl_nadam_hyper_params* h, vcl_stream_t stream) {
    VCL_REQUIRE(p && h && p->param && p->grad && h->m && h->v && p->size >= 0);
    if (p->size == 0) return VCL_OK;
    NadamArgs a = {h->mu, h->mu0, h->mu1, h->rho, h->alpha, h->mut0, h->mut1, h->rhot, h->eps};
    LAUNCH(k_nadam, grid_for((size_t)p->size), BS, as_hip(stream), (size_t)p->size, p->param, p->grad, h->m, h->v, a);
    return VCL_OK;
}

extern "C" int vcl_optimize_random_params(const vcl_optimize_params* p, uint32_t seed, float lo, float hi, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->param && p->size >= 0);
    if (p->size == 0) return VCL_OK;
    LAUNCH(k_random_fill, grid_for((size_t)p->size), BS, as_hip(stream), (size_t)p->size, p->param, (u32)seed, 0xdeadbeefu, lo, hi, 0);
    return VCL_OK;
}

extern "C" int vcl_optimize_clamp_params(const vcl_optimize_params* p, float lo, float hi, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->param && p->size >= 0);
    if (p->size == 0) return VCL_OK;
    LAUNCH(k_clamp_buf, grid_for((size_t)p->size), BS, as_hip(stream), (size_t)p->size, p->param, p->param, lo, hi);
    return VCL_OK;
}

extern "C" int vcl_buffer_linear(float* dst, const float* src, size_t n, float w, float b, vcl_stream_t stream) {
    VCL_REQUIRE(dst && src);
    if (n == 0) return VCL_OK;
    LAUNCH(k_linear, grid_for(n), BS, as_hip(stream), n, dst, src, w, b);
    return VCL_OK;
}

extern "C" int vcl_buffer_add_random(float* dst, size_t n, float lo, float hi, uint32_t seed, vcl_stream_t stream) {
    VCL_REQUIRE(dst);
    if (n == 0) return VCL_OK;
    LAUNCH(k_random_fill, grid_for(n), BS, as_hip(stream), n, dst, (u32)seed, 0xba5ec0deu, lo, hi, 1);
    return VCL_OK;
}

extern "C" int vcl_buffer_clamp(float* dst, const float* src, size_t n, float lo, float hi, vcl_stream_t stream) {
    VCL_REQUIRE(dst && src);
    if (n == 0) return VCL_OK;
    LAUNCH(k_clamp_buf, grid_for(n), BS, as_hip(stream), n, dst, src, lo, hi);
    return VCL_OK;
}

extern "C" int vcl_buffer_step(float* dst, const float* src, size_t n, float thr, vcl_stream_t stream) {
    VCL_REQUIRE(dst && src);
    if (n == 0) return VCL_OK;
    LAUNCH(k_step_buf, grid_for(n), BS, as_hip(stream), n, dst, src, thr);
    return VCL_OK;
}

extern "C" int vcl_buffer_normalize(float* dst, const float* src, int rows, int dim, vcl_stream_t stream) {
    VCL_REQUIRE(dst && src && rows >= 0 && dim > 0);
    if (rows == 0) return VCL_OK;
    LAUNCH(k_normalize_rows, cdiv((size_t)rows, BS), BS, as_hip(stream), rows, dim, dst, src);
    return VCL_OK;
}

extern "C" int vcl_buffer_set_value(float* dst, int rows, int dim, const float* value, vcl_stream_t stream) {
    VCL_REQUIRE(dst && value && rows >= 0 && dim > 0 && dim <= 16);
    if (rows == 0) return VCL_OK;
    Value16 v;
    std::memset(&v, 0, sizeof(v));
    std::memcpy(v.v, value, sizeof(float) * dim);
    LAUNCH(k_set_value, cdiv((size_t)rows, BS), BS, as_hip(stream), rows, dim, dst, v);
    return VCL_OK;
}

extern "C" int vcl_buffer_spherical_random(float* dst, int count, uint32_t seed, vcl_stream_t stream) {
    VCL_REQUIRE(dst && count >= 0);
    if (count == 0) return VCL_OK;
    LAUNCH(k_spherical_random, cdiv((size_t)count, BS), BS, as_hip(stream), count, dst, (u32)seed);
    return VCL_OK;
}

extern "C" int vcl_attribute_distance_error(const float* pred, const float* target, int num, int dim, float* out, vcl_stream_t stream) {
    VCL_REQUIRE(pred && target && out && num > 0 && dim > 0);
    LAUNCH(k_distance_error, 1, 1024, as_hip(stream), num, dim, pred, target, out);
    return VCL_OK;
}

extern "C" int vcl_texture_bilinear_downsampling(const vcl_texture* t, vcl_stream_t stream) {
    VCL_REQUIRE(t && t->width > 0 && t->height > 0 && t->channel > 0 && t->miplevel >= 1 && t->miplevel <= VCL_TEX_MAX_MIP_LEVEL);
    int w = t->width, h = t->height;
    for (int l = 1; l < t->miplevel; l++) {
        VCL_REQUIRE(t->texture[l - 1] && t->texture[l]);
        w >>= 1; h >>= 1;
        if (w * h == 0) break;
        LAUNCH(k_bilinear_down, cdiv((size_t)w * h, BS), BS, as_hip(stream), (const float*)t->texture[l - 1], t->texture[l], w, h, t->channel);
    }
    return VCL_OK;
}

extern "C" int vcl_sg_bake(const float* axis, const float* sharp, const float* amp, int num, int channel, const vcl_texture* t,
                           vcl_stream_t stream) {
    VCL_REQUIRE(axis && sharp && amp && t && t->texture[0] && num >= 0 && channel > 0 && t->channel == channel && t->width > 0 && t->height > 0);
    int w = t->width, h = t->height;
    for (int l = 0; l < t->miplevel; l++, w >>= 1, h >>= 1) {
        VCL_REQUIRE(t->texture[l]);
        size_t n = (size_t)w * h * channel;
        if (n) LAUNCH(k_linear, grid_for(n), BS, as_hip(stream), n, t->texture[l], (const float*)t->texture[l], 0.f, 0.f);
    }
    if (num == 0) return VCL_OK;
    LAUNCH(k_sg_bake, cdiv((size_t)t->width * t->height, BS), BS, as_hip(stream), num, channel, axis, sharp, amp, t->width, t->height, t->texture[0]);
    return VCL_OK;
}

extern "C" int vcl_sg_normalize(float* axis, float* sharp, float* amp, int num, int channel, vcl_stream_t stream) {
    VCL_REQUIRE(axis && sharp && amp && num >= 0 && channel > 0);
    if (num == 0) return VCL_OK;
    LAUNCH(k_sg_normalize, cdiv((size_t)num, BS), BS, as_hip(stream), num, channel, axis, sharp, amp);
    return VCL_OK;
}

extern "C" int vcl_sg_disperse(float* axis, const float* sharp, const float* amp, int num, int channel, float* buffer, vcl_stream_t stream) {
    VCL_REQUIRE(axis && sharp && amp && buffer && num >= 0 && channel > 0);
    if (num == 0) return VCL_OK;
    LAUNCH(k_sg_disperse, cdiv((size_t)num, BS), BS, as_hip(stream), num, channel, (const float*)axis, sharp, amp, buffer);
    LAUNCH(k_add3, cdiv((size_t)num * 3, BS), BS, as_hip(stream), num, axis, (const float*)buffer);
    return VCL_OK;
}

// texture loss workspace: the per-element map (size floats), then one partial per block
extern "C" size_t vcl_texture_loss_workspace_size(int size) {
    size_t n = size > 0 ? (size_t)size : 1;
    return ((n * sizeof(float) + 255) & ~(size_t)255) + sizeof(float) * cdiv(n, BS) + 256;
}

extern "C" int vcl_mse_texture_loss(const vcl_loss_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->predict && p->target && p->buffer && p->loss && p->width > 0 && p->height > 0 && p->depth > 0);
    VCL_REQUIRE((size_t)p->size == (size_t)p->width * p->height * p->depth);
    size_t n = (size_t)p->size;
    float* part = (float*)((char*)p->buffer + ((n * sizeof(float) + 255) & ~(size_t)255));
    unsigned nb = cdiv(n, BS);
    LAUNCH(k_texture_loss, nb, BS, as_hip(stream), p->width, p->height, p->depth, p->predict, p->target, p->buffer, part);
    LAUNCH(k_sum_partials, 1, 1024, as_hip(stream), (const float*)part, (size_t)nb, 1.0 / (double)(p->size / 2), p->loss);
    return VCL_OK;
}

// SSIM workspace: the 1 - S map, three coefficient maps (exact gradient), block partials
static size_t ssim_blocks(int W, int H) { return (size_t)cdiv((size_t)W, SS_T) * cdiv((size_t)H, SS_T); }
static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" size_t vcl_ssim_workspace_size(int W, int H, int D) {
    if (W <= 0 || H <= 0 || D <= 0) return 0;
    size_t N = (size_t)W * H * D;
    return 4 * al256(N * sizeof(float)) + al256(ssim_blocks(W, H) * sizeof(float)) + 256;
}

extern "C" int vcl_ssim_loss_backward(const vcl_loss_params* p, int mode, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->predict && p->target && p->grad && p->buffer && p->loss && p->width > 0 && p->height > 0 && p->depth > 0);
    VCL_REQUIRE(mode == VCL_SSIM_GRAD_REF || mode == VCL_SSIM_GRAD_EXACT);
    VCL_REQUIRE((size_t)p->size == (size_t)p->width * p->height * p->depth);
    const size_t N = (size_t)p->size, nb = ssim_blocks(p->width, p->height);
    char* ws = (char*)p->buffer;
    SsimArgs a;
    a.W = p->width; a.H = p->height; a.D = p->depth; a.mode = mode;
    a.pred = p->predict; a.target = p->target; a.grad = p->grad;
    a.map = (float*)ws;
    a.alpha = (float*)(ws + al256(N * 4));
    a.beta = (float*)(ws + 2 * al256(N * 4));
    a.gam = (float*)(ws + 3 * al256(N * 4));
    a.part = (float*)(ws + 4 * al256(N * 4));
    LAUNCH(k_ssim_stats, (unsigned)nb, BS, as_hip(stream), a);
    if (mode == VCL_SSIM_GRAD_EXACT) LAUNCH(k_ssim_grad, (unsigned)nb, BS, as_hip(stream), a);
    LAUNCH(k_sum_partials, 1, 1024, as_hip(stream), (const float*)a.part, nb, 1.0 / (4.0 * (double)N), p->loss);
    return VCL_OK;
}

extern "C" int vcl_filter_forward(const vcl_filter_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->sigma && p->filter && p->in && p->buf && p->out && p->width > 0 && p->height > 0 && p->depth > 0 && p->channel > 0);
    hipStream_t s = as_hip(stream);
    FilterArgs f = {p->width, p->height, p->depth, p->channel};
    size_t P = (size_t)p->width * p->height * p->depth;
    LAUNCH(k_filter_weights, 1, 64, s, p->sigma, p->filter, (float*)nullptr);
    LAUNCH((k_filter_pass<0, 1>), grid_for(P), BS, s, f, (const float*)p->filter, (const float*)nullptr, p->in, p->buf, (float*)nullptr);
    LAUNCH((k_filter_pass<1, 1>), grid_for(P), BS, s, f, (const float*)p->filter, (const float*)nullptr, (const float*)p->buf, p->out, (float*)nullptr);
    return VCL_OK;
}

extern "C" int vcl_filter_backward(const vcl_filter_params* p, const vcl_filter_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->sigma && p->filter && p->in && g->in && g->sigma && g->filter && g->buf && g->bufx && g->bufy && g->out);
    VCL_REQUIRE(p->width > 0 && p->height > 0 && p->depth > 0 && p->channel > 0);
    hipStream_t s = as_hip(stream);
    FilterArgs f = {p->width, p->height, p->depth, p->channel};
    size_t P = (size_t)p->width * p->height * p->depth;
    unsigned nb = cdiv(P, BS);
    LAUNCH(k_filter_weights, 1, 64, s, p->sigma, p->filter, g->filter);
    LAUNCH((k_filter_pass<0, 2>), grid_for(P), BS, s, f, (const float*)p->filter, (const float*)g->filter, g->out, g->buf, g->bufy);
    LAUNCH(k_filter_bwd_v, nb, BS, s, f, (const float*)p->filter, (const float*)g->filter, (const float*)g->buf, (const float*)g->bufy, p->in,
           g->in, g->bufx);
    LAUNCH(k_filter_sigma, 1, 1024, s, (const float*)g->bufx, (size_t)nb, p->sigma, g->sigma);
    return VCL_OK;
}

extern "C" int vcl_compute_normal_forward(const vcl_compute_normal_params* p, vcl_stream_t stream) {
    VCL_REQUIRE(p && p->pos && p->vao && p->out && p->sum && p->posNum > 0 && p->vaoNum >= 0);
    VCL_REQUIRE(p->angleMode == VCL_NORMAL_ANGLE_REF || p->angleMode == VCL_NORMAL_ANGLE_EXACT);
    hipStream_t s = as_hip(stream);
    size_t n = (size_t)p->posNum * 3;
    LAUNCH(k_zero, grid_for(n), BS, s, n, p->sum);
    if (p->vaoNum) LAUNCH(k_cn_scatter, cdiv((size_t)p->vaoNum, BS), BS, s, p->vaoNum, p->pos, p->vao, p->angleMode, p->sum);
    LAUNCH(k_cn_normalize, cdiv((size_t)p->posNum, BS), BS, s, p->posNum, (const float*)p->sum, p->out);
    return VCL_OK;
}

extern "C" int vcl_compute_normal_backward(const vcl_compute_normal_params* p, const vcl_compute_normal_grad_params* g, vcl_stream_t stream) {
    VCL_REQUIRE(p && g && p->pos && p->vao && p->out && p->sum && g->out && p->posNum > 0 && p->vaoNum >= 0);
    VCL_REQUIRE(p->angleMode == VCL_NORMAL_ANGLE_REF || p->angleMode == VCL_NORMAL_ANGLE_EXACT);
    if (!g->pos || !p->vaoNum) return VCL_OK;
    LAUNCH(k_cn_backward, cdiv((size_t)p->vaoNum, BS), BS, as_hip(stream), p->vaoNum, p->pos, p->vao, p->angleMode, (const float*)p->sum,
           (const float*)p->out, g->out, g->pos);
    return VCL_OK;
}

extern "C" int vcl_rasterize_wireframe(int W, int H, int D, const float* rast, const float* db, float* out, vcl_stream_t stream) {
    VCL_REQUIRE(rast && db && out && W > 0 && H > 0 && D > 0);
    size_t P = (size_t)W * H * D;
    LAUNCH(k_wireframe, cdiv(P, BS), BS, as_hip(stream), P, rast, db, out);
    return VCL_OK;
}

extern "C" int vcl_rasterize_idhash(int W, int H, int D, const float* rast, float* out, vcl_stream_t stream) {
    VCL_REQUIRE(rast && out && W > 0 && H > 0 && D > 0);
    size_t P = (size_t)W * H * D;
    LAUNCH(k_idhash, cdiv(P, BS), BS, as_hip(stream), P, rast, out);
    return VCL_OK;
}
