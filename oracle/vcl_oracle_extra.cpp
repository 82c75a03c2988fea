// vcl_oracle_extra.cpp -- CPU restatement of the reference ops beyond the fit chain: noise,
// Gaussian filter, SSIM and texture losses, SGD / Nadam and the optimizer helpers, buffer
// utilities (linear, random, clamp, step, normalize, set value, spherical random, distance
// error), texture bilinear downsampling, SG bake / normalize / disperse, ComputeNormal, and the
// debug raster shadings (wireframe, id hash).
//
// TEST INFRASTRUCTURE ONLY, like vcl_oracle.cpp (see its header): the GPU library is compared
// against these functions; the product path never calls them.  Templated on T (float follows
// the reference's arithmetic, double is the FP64 yardstick) where the op is continuous.
// Citations name files of the reference (236088/vcl, nvdiffrast/nvdiffrast/<file>:<line>).
// Parity status per function:
//   pinned          -- checked in tests/test_oracle_extra.py (test named in the comment);
//   parity unpinned -- no reference test, fixture or published vector fixes it (named below).
#include "oracle_common.hpp"

// ------------------------------------------------------------------------------------
// getUniform (common.h:194-207): Bob Jenkins' 96-bit mix (lookup2.c, 1996, "mix()"), then the
// low 23 bits of c scaled to [0, 1].
// parity unpinned for the exact bits: no reference test or published vector of this mix is
// available offline; tests/test_oracle_extra.py::test_uniform_distribution checks the range,
// the mean and variance of U[0,1] and decorrelation between seeds and neighbouring indices.
static inline void jenkins_mix(u32& a, u32& b, u32& c) {
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
static inline float uniform01(u32 a, u32 b, u32 c) {
    jenkins_mix(a, b, c);
    const int d = 0x007fffff;
    return float(c & d) / float(d);
}
extern "C" void orc_uniform(size_t n, const u32* a, u32 b, u32 c, float* out) {
    for (size_t i = 0; i < n; i++) out[i] = uniform01(a[i], b, c);
}

static inline float clampf(float a, float lo, float hi) { return a < lo ? lo : a > hi ? hi : a; }

// ------------------------------------------------------------------------------------
// Noise::forward (extention.cu:13-30): out = in + intensity * (U(pidx, seed + i, 0xcafef00d)
// - in) per channel i.  Every layer is processed (the reference's 2D launch covers layer 0
// only, extention.cu:28; layers above 0 are broken upstream anyway, SURVEY A-19).
// pinned: intensity 0 is the identity, intensity 1 gives U (test_noise_closed_forms).
extern "C" void orc_noise(int W, int H, int depth, int C, float intensity, const float* in, float* out, u32 seed) {
    size_t P = (size_t)W * H * depth;
    for (size_t p = 0; p < P; p++)
        for (int i = 0; i < C; i++) {
            float x = in[p * C + i];
            out[p * C + i] = x + intensity * (uniform01((u32)p, seed + (u32)i, 0xcafef00du) - x);
        }
}

// ------------------------------------------------------------------------------------
// Optimizers (optimize.cu).  All flat over `size` elements (the reference's 3D grid over
// width x height x depth = size).
// SGD (optimize.cu:64-82): AddNaNcheck(param, (float)(-eta * grad)).
// pinned: textbook update (test_sgd_nadam_recurrences).
extern "C" void orc_sgd(size_t n, float* param, const float* grad, double eta) {
    for (size_t i = 0; i < n; i++) add_nan_check(param[i], (float)(-eta * grad[i]));
}

// Nadam kernel (optimize.cu:143-155) with the host-advanced state of Nadam::setHyperParams
// and Nadam::step (:126-141, :157-167): m, v updated in double and stored as float, then
// mhat = mu1 / (1 - mut1) * m + (1 - mu0) / (1 - mut0) * g, vhat = v / (1 - rhot).
// pinned: Dozat's Nadam recurrence with the reference's schedule mu_t = mu (1 - 0.96^(0.004 t)/2)
// (test_sgd_nadam_recurrences).
extern "C" void orc_nadam(size_t n, float* param, const float* grad, float* m, float* v, double mu, double mu0, double mu1, double rho,
                          double alpha, double mut0, double mut1, double rhot, double eps) {
    for (size_t i = 0; i < n; i++) {
        double g = grad[i];
        m[i] = (float)(mu * m[i] + (1 - mu) * g);
        v[i] = (float)(rho * v[i] + (1 - rho) * g * g);
        double mh = mu1 / (1 - mut1) * m[i] + (1 - mu0) / (1 - mut0) * g;
        double vh = 1 / (1 - rhot) * v[i];
        add_nan_check(param[i], (float)(-mh * alpha / (std::sqrt(vh) + eps)));
    }
}

// Optimizer::randomParams (optimize.cu:30-44): param = min + (max - min) U(i, seed, 0xdeadbeef);
// clampParams (:46-60) is orc_clamp_buf.
// pinned: range and distribution (test_buffer_utilities).
extern "C" void orc_random_params(size_t n, float* param, u32 seed, float lo, float hi) {
    for (size_t i = 0; i < n; i++) param[i] = lo + (hi - lo) * uniform01((u32)i, seed, 0xdeadbeefu);
}

// ------------------------------------------------------------------------------------
// Buffer utilities (buffer.cu:3-110, used by Buffer / Attribute / Texture / SGBuffer).
// pinned: closed forms (test_buffer_utilities).
extern "C" void orc_linear(size_t n, float* dst, const float* src, float w, float b) {  // linerKernel
    for (size_t i = 0; i < n; i++) dst[i] = src[i] * w + b;
}
extern "C" void orc_add_random(size_t n, float* dst, float lo, float hi, u32 seed) {  // randomKernel
    for (size_t i = 0; i < n; i++) dst[i] += lo + (hi - lo) * uniform01((u32)i, seed, 0xba5ec0deu);
}
extern "C" void orc_clamp_buf(size_t n, float* dst, const float* src, float lo, float hi) {  // clampKernel
    for (size_t i = 0; i < n; i++) dst[i] = clampf(src[i], lo, hi);
}
extern "C" void orc_step_buf(size_t n, float* dst, const float* src, float thr) {  // stepKernel
    for (size_t i = 0; i < n; i++) dst[i] = src[i] < thr ? 0.f : 1.f;
}
// normalizeKernel (buffer.cu:83-97): rows of `dim` values scaled by 1/|row| in place (dst *= il);
// zero rows are left as they are.
extern "C" void orc_normalize_rows(int rows, int dim, float* dst, const float* src) {
    for (int r = 0; r < rows; r++) {
        float l2 = 0.f;
        for (int i = 0; i < dim; i++) l2 += src[(size_t)r * dim + i] * src[(size_t)r * dim + i];
        if (l2 <= 0) continue;
        float il = 1.f / std::sqrt(l2);
        for (int i = 0; i < dim; i++) dst[(size_t)r * dim + i] *= il;
    }
}
// sphericalRandomKernel (buffer.cu:23-36): uniform directions from z in [-1, 1], theta in [-pi, pi].
extern "C" void orc_spherical_random(int n, float* dst, u32 seed) {
    for (int i = 0; i < n; i++) {
        float theta = (uniform01((u32)i, seed, 0xba5eba11u) * 2.f - 1.f) * 3.14159265f;
        float z = uniform01((u32)i, seed, 0xdeadba11u) * 2.f - 1.f;
        float c = std::sqrt(1.f - z * z);
        dst[3 * i] = c * std::cos(theta);
        dst[3 * i + 1] = c * std::sin(theta);
        dst[3 * i + 2] = z;
    }
}
// Attribute::distanceError (buffer.cu:321-344): mean over vertices of |predict - target|
// (the reference sums with float atomics; here in double).
extern "C" double orc_distance_error(int num, int dim, const float* pred, const float* target) {
    double s = 0;
    for (int i = 0; i < num; i++) {
        float d = 0.f;
        for (int k = 0; k < dim; k++) { float e = pred[(size_t)i * dim + k] - target[(size_t)i * dim + k]; d += e * e; }
        s += std::sqrt(d);
    }
    return s / num;
}

// Texture::bilinearDownsampling (buffer.cu:410-447): level l from level l-1 with the separable
// [1 3 3 1] / 8 filter centred on each 2 x 2 block, taps outside the level dropped (not
// renormalised), channels 0..min(C, 3)-1 (the reference writes three channels).
// pinned: separable correlation by scipy.ndimage (test_bilinear_downsampling).
extern "C" void orc_bilinear_downsampling(int tw, int th, int C, int L, float* levels) {
    static const float filt[4] = {.125f, .375f, .375f, .125f};
    size_t off = 0;
    int w = tw, h = th;
    for (int l = 1; l < L; l++) {
        const float* src = levels + off;
        off += (size_t)w * h * C;
        float* dst = levels + off;
        w >>= 1; h >>= 1;
        int nc = C < 3 ? C : 3;
        for (int py = 0; py < h; py++)
            for (int px = 0; px < w; px++) {
                int x2 = px << 1, y2 = py << 1, W2 = w << 1, H2 = h << 1;
                int xs = x2 > 0 ? -1 : 0, xe = x2 < W2 - 2 ? 2 : 1;
                int ys = y2 > 0 ? -1 : 0, ye = y2 < H2 - 2 ? 2 : 1;
                size_t p = (size_t)px + (size_t)w * py;
                for (int i = 0; i < nc; i++) dst[p * C + i] = 0.f;
                for (int x = xs; x <= xe; x++)
                    for (int y = ys; y <= ye; y++) {
                        float f = filt[x + 1] * filt[y + 1];
                        size_t q = (size_t)(x2 + x) + (size_t)W2 * (y2 + y);  // the reference strides by 2w (buffer.cu:429)
                        for (int i = 0; i < nc; i++) dst[p * C + i] += src[q * C + i] * f;
                    }
            }
    }
}

// SGBuffer::bake (buffer.cu:734-761) into level 0 (the caller zeroes the texture first,
// Texture::liner(texture, 0, 0)): equirectangular directions
// phi = (1 - 2 px / W) pi, theta = (2 py / H - 1) pi / 2, axis = (sin phi cos theta, sin theta,
// cos phi cos theta); tex += amplitude_i exp(sharpness_i (axis . mu_i - 1)) in lobe order with
// AddNaNcheck.
// pinned: a lobe aimed at a texel centre gives its amplitude there and decays with angle
// (test_sg_bake_normalize_disperse).
extern "C" void orc_sg_bake(int num, int C, const float* axis, const float* sharp, const float* amp, int W, int H, float* tex0) {
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            size_t p = (size_t)px + (size_t)W * py;
            float phi = (1.f - 2.f * (float)px / (float)W) * 3.14159265f;
            float theta = (2.f * (float)py / (float)H - 1.f) * 1.5707963f;
            float ax = std::sin(phi) * std::cos(theta), ay = std::sin(theta), az = std::cos(phi) * std::cos(theta);
            for (int i = 0; i < num; i++) {
                float d = ax * axis[3 * i] + ay * axis[3 * i + 1] + az * axis[3 * i + 2];
                float v = std::exp(sharp[i] * (d - 1.f));
                for (int k = 0; k < C; k++) add_nan_check(tex0[p * C + k], amp[(size_t)i * C + k] * v);
            }
        }
}

// SGBuffer::normalize (buffer.cu:685-704): unit axes (rows of 3: the reference strides by
// `channel`, equal for channel = 3 and out of bounds otherwise), sharpness clamped to [0, 8],
// amplitude to [0, 1].
extern "C" void orc_sg_normalize(int num, int C, float* axis, float* sharp, float* amp) {
    orc_normalize_rows(num, 3, axis, axis);
    orc_clamp_buf((size_t)num, sharp, sharp, 0.f, 8.f);
    orc_clamp_buf((size_t)num * C, amp, amp, 0.f, 1.f);
}

// SGBufferGrad::disperse (buffer.cu:771-804): each lobe's axis moves by the sum over lobes of
// the unit separation times 1 / max(|separation|^2 sharpness |amplitude difference|^2, 1e-6).
// The reference reads the amplitude difference from sharpness[i * channel + k] (out of bounds
// for i >= num / channel); the amplitudes are the evident operand and are used here.
extern "C" void orc_sg_disperse(int num, int C, float* axis, const float* sharp, const float* amp, float* buffer) {
    for (int p = 0; p < num; p++) {
        float a0 = axis[3 * p], a1 = axis[3 * p + 1], a2 = axis[3 * p + 2], s = sharp[p];
        float f0 = 0.f, f1 = 0.f, f2 = 0.f;
        for (int i = 0; i < num; i++) {
            float sq = 0.f;
            for (int k = 0; k < C; k++) { float d = amp[(size_t)p * C + k] - amp[(size_t)i * C + k]; sq += d * d; }
            float v0 = a0 - axis[3 * i], v1 = a1 - axis[3 * i + 1], v2 = a2 - axis[3 * i + 2];
            float l2 = (float)std::max((double)(v0 * v0 + v1 * v1 + v2 * v2), 1e-6);
            float il = 1.f / std::sqrt(l2);
            v0 *= il; v1 *= il; v2 *= il;
            float w = (float)(1.0 / std::max((double)(l2 * s * sq), 1e-6));
            f0 += v0 * w; f1 += v1 * w; f2 += v2 * w;
        }
        buffer[3 * p] = f0; buffer[3 * p + 1] = f1; buffer[3 * p + 2] = f2;
    }
    for (int p = 0; p < num; p++)
        for (int k = 0; k < 3; k++) axis[3 * p + k] += buffer[3 * p + k];
}

// ------------------------------------------------------------------------------------
// MSELoss::textureloss (loss.cu:59-91): per element (HWC with `depth` channels,
// pidx = (px + W py) depth + pz) the squared difference, zero where
// (px < W/4 or W/2 <= px) and (py < W/4 or W/2 <= py) (both tests use the width);
// loss = sum / (size / 2) with integer division.  No gradient is written.
// pinned: closed form (test_texture_loss).
extern "C" double orc_texture_loss(int W, int H, int depth, const float* pred, const float* target, float* buffer) {
    int size = W * H * depth;
    int w = W / 4, w2 = W / 2;
    double s = 0;
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++)
            for (int pz = 0; pz < depth; pz++) {
                size_t p = ((size_t)px + (size_t)W * py) * depth + pz;
                float v = 0.f;
                if (!((px < w || w2 <= px) && (py < w || w2 <= py))) { float t = pred[p] - target[p]; v = t * t; }
                buffer[p] = v;
                s += v;
            }
    return s / (double)(size / 2);
}

// ------------------------------------------------------------------------------------
// SSIMLoss::backward (loss.cu:93-174).  Per element (HWC with `depth` channels) the window is
// the pixel's (2k+1)^2 neighbourhood with k = 5 clipped to the image, and every mean and
// (co)variance divides by k^2 = 25, not by the window's size (loss.cu:155).  Sums run over
// ix (outer) then iy (inner), means first, then the centred second moments.
//   S = (2 mx my + c1)(2 cov + c2) / ((mx^2 + my^2 + c1)(vx + vy + c2)), c1 = 1e-4, c2 = 1e-3;
//   buffer = 1 - S; loss = sum(1 - S) / (4 size) (loss.cu:173 divides by the byte size).
// Gradient modes:
//   ref (0)  : loss.cu:146-149 as written -- the centre pixel's term of its own window only,
//              with (1 - S) where the derivative has S;
//   exact (1): d/d predict of sum_p (1 - S_p) over every window that holds the element.
// pinned: S by scipy box sums, the ref gradient on a 5 x 5 image (where the window is the
// whole image and 1/25 is its mean) by finite differences of the numerator and denominator
// products, the exact gradient by finite differences of sum(1 - S) (test_ssim_*).
template <class T>
struct SsimStats { T mx, my, vx, vy, cov; int n; T sx, sy; };

template <class T>
static SsimStats<T> ssim_window(int W, int H, int D, const T* x, const T* y, int px, int py, int pz, int k, T inv) {
    int sx = std::max(px - k, 0), ex = std::min(px + k, W - 1);
    int sy = std::max(py - k, 0), ey = std::min(py + k, H - 1);
    SsimStats<T> s;
    T ax = 0, ay = 0;
    for (int ix = sx; ix <= ex; ix++)
        for (int iy = sy; iy <= ey; iy++) {
            size_t q = ((size_t)ix + (size_t)W * iy) * D + pz;
            ax += x[q]; ay += y[q];
        }
    s.sx = ax; s.sy = ay;
    s.mx = ax * inv; s.my = ay * inv;
    T vx = 0, vy = 0, cv = 0;
    for (int ix = sx; ix <= ex; ix++)
        for (int iy = sy; iy <= ey; iy++) {
            size_t q = ((size_t)ix + (size_t)W * iy) * D + pz;
            T dx = x[q] - s.mx, dy = y[q] - s.my;
            vx += dx * dx; vy += dy * dy; cv += dx * dy;
        }
    s.vx = vx * inv; s.vy = vy * inv; s.cov = cv * inv;
    s.n = (ex - sx + 1) * (ey - sy + 1);
    return s;
}

template <class T>
static double ssim(int W, int H, int D, const T* pred, const T* target, T* grad, T* buffer, int mode) {
    const int k = 5;
    const T c1 = (T)1e-4f, c2 = (T)1e-3f, inv = T(1) / (T)(k * k);
    size_t N = (size_t)W * H * D;
    std::vector<T> alpha, beta, gam;
    if (mode == 1) { alpha.assign(N, 0); beta.assign(N, 0); gam.assign(N, 0); }
    parallel_for(H, [&](int r0, int r1, int) {
        for (int py = r0; py < r1; py++)
            for (int px = 0; px < W; px++)
                for (int pz = 0; pz < D; pz++) {
                    size_t p = ((size_t)px + (size_t)W * py) * D + pz;
                    SsimStats<T> s = ssim_window<T>(W, H, D, pred, target, px, py, pz, k, inv);
                    T n1 = T(2) * s.mx * s.my + c1, d1 = s.mx * s.mx + s.my * s.my + c1;
                    T n2 = T(2) * s.cov + c2, d2 = s.vx + s.vy + c2;
                    T dinv = T(1) / (d1 * d2);
                    T S = n1 * n2 * dinv;
                    buffer[p] = T(1) - S;
                    if (mode == 0) {
                        if (grad)
                            grad[p] = -(s.my * n2 + n1 * (target[p] - s.my) - buffer[p] * (s.mx * d2 + d1 * (pred[p] - s.mx))) * T(2) * inv * dinv;
                    } else {
                        // dS/dx_q = inv [A - 2B (mx + ex) - C (my + ey)] + inv [2B x_q + C y_q], with
                        // A = dS/dmx, B = dS/dvx, C = dS/dcov and e = inv (sum - n m)
                        T A = T(2) * (s.my * n2 * dinv - S * s.mx / d1);
                        T B = -S / d2;
                        T Cc = T(2) * n1 * dinv;
                        T ex = inv * (s.sx - (T)s.n * s.mx), ey = inv * (s.sy - (T)s.n * s.my);
                        alpha[p] = A - T(2) * B * (s.mx + ex) - Cc * (s.my + ey);
                        beta[p] = T(2) * B;
                        gam[p] = Cc;
                    }
                }
    });
    if (mode == 1 && grad) {
        // grad_q = -inv sum over windows p holding q of (alpha_p + beta_p x_q + gamma_p y_q)
        parallel_for(H, [&](int r0, int r1, int) {
            for (int py = r0; py < r1; py++)
                for (int px = 0; px < W; px++)
                    for (int pz = 0; pz < D; pz++) {
                        size_t q = ((size_t)px + (size_t)W * py) * D + pz;
                        int sx = std::max(px - k, 0), ex = std::min(px + k, W - 1);
                        int sy = std::max(py - k, 0), ey = std::min(py + k, H - 1);
                        T a = 0, b = 0, g = 0;
                        for (int ix = sx; ix <= ex; ix++)
                            for (int iy = sy; iy <= ey; iy++) {
                                size_t p = ((size_t)ix + (size_t)W * iy) * D + pz;
                                a += alpha[p]; b += beta[p]; g += gam[p];
                            }
                        grad[q] = -inv * (a + b * pred[q] + g * target[q]);
                    }
        });
    }
    double sum = 0;
    for (size_t p = 0; p < N; p++) sum += (double)buffer[p];
    return sum / (4.0 * (double)N);
}

// ------------------------------------------------------------------------------------
// Filter (filter.cu:1-186): separable Gaussian blur, horizontal then vertical, each output the
// sum over the (2k+1) taps inside the image (weights not renormalised at the borders),
// weights w_i = exp(-i^2 / (2 sigma^2)) / sum_j exp(-j^2 / (2 sigma^2)) for |i| <= k,
// k = min(ceil(4 sigma), 127) (FILTER_MAX_SIZE 255, filter.cu:2,62-80).  HWC with `channel`
// values per pixel.  Layers read their own layer (the reference reads layer 0 for every layer,
// filter.cu:40,56) and the vertical pass strides by the width (the reference uses the height,
// filter.cu:56; equal for square images).
// Backward: grad.in = the same two passes applied to grad.out (the adjoint of a symmetric
// truncated correlation is itself); d loss / d sigma exactly:
//   sum_p in_p [(H_f V_g + H_g V_f) gout - 2 m2 H_f V_f gout]_p / sigma^3,
// with g_i = w_i i^2 and m2 = sum_i w_i i^2 (the derivative of the normalised weights).  The
// reference's FilterBackwardKernelFinal adds each pixel's term once per tap of its vertical
// window and uses 2 / sigma for 2 m2 / sigma^3 (filter.cu:141-157,182).
// pinned: forward by scipy.ndimage.correlate1d, input grad by the adjoint identity, sigma grad
// by central differences (test_filter_*).
template <class T>
static int filter_weights(T sigma, std::vector<T>& w, T& m2) {
    int k = (int)std::ceil((double)sigma * 4.0);
    if (k * 2 > 255) k = 255 / 2;
    if (k < 0) k = 0;
    w.assign(2 * k + 1, T(0));
    T s2 = T(0.5) / (sigma * sigma), s = 0;
    for (int i = -k; i <= k; i++) { w[i + k] = std::exp(-(T)(i * i) * s2); s += w[i + k]; }
    s = T(1) / s;
    m2 = 0;
    for (int i = -k; i <= k; i++) { w[i + k] *= s; m2 += w[i + k] * (T)(i * i); }
    return k;
}

// one pass along x (axis 0) or y (axis 1) with weights w (2k+1 taps)
template <class T>
static void filter_pass(int W, int H, int D, int C, int k, const std::vector<T>& w, int axis, const T* in, T* out) {
    parallel_for(D * H, [&](int r0, int r1, int) {
        for (int row = r0; row < r1; row++) {
            int pz = row / H, py = row % H;
            for (int px = 0; px < W; px++) {
                int pos = axis == 0 ? px : py, len = axis == 0 ? W : H;
                int s = pos - k, f = 0;
                if (s < 0) { f = -s; s = 0; }
                int e = std::min(pos + k, len - 1);
                size_t p = (size_t)px + (size_t)W * (py + (size_t)H * pz);
                for (int c = 0; c < C; c++) out[p * C + c] = 0;
                for (int i = s, fi = f; i <= e; i++, fi++) {
                    size_t q = axis == 0 ? (size_t)i + (size_t)W * (py + (size_t)H * pz) : (size_t)px + (size_t)W * (i + (size_t)H * pz);
                    for (int c = 0; c < C; c++) out[p * C + c] += in[q * C + c] * w[fi];
                }
            }
        }
    });
}

template <class T>
static void filter_fwd(int W, int H, int D, int C, T sigma, const T* in, T* out) {
    std::vector<T> w; T m2;
    int k = filter_weights<T>(sigma, w, m2);
    std::vector<T> buf((size_t)W * H * D * C);
    filter_pass<T>(W, H, D, C, k, w, 0, in, buf.data());
    filter_pass<T>(W, H, D, C, k, w, 1, buf.data(), out);
}

template <class T>
static double filter_bwd(int W, int H, int D, int C, T sigma, const T* in, const T* gout, T* gin) {
    std::vector<T> w, g; T m2;
    int k = filter_weights<T>(sigma, w, m2);
    g.resize(w.size());
    for (int i = -k; i <= k; i++) g[i + k] = w[i + k] * (T)(i * i);
    size_t n = (size_t)W * H * D * C;
    std::vector<T> hf(n), hg(n), vg(n), vfg(n);
    filter_pass<T>(W, H, D, C, k, w, 0, gout, hf.data());
    filter_pass<T>(W, H, D, C, k, g, 0, gout, hg.data());
    filter_pass<T>(W, H, D, C, k, w, 1, hf.data(), gin);
    filter_pass<T>(W, H, D, C, k, g, 1, hf.data(), vg.data());
    filter_pass<T>(W, H, D, C, k, w, 1, hg.data(), vfg.data());
    double s = 0;
    for (size_t i = 0; i < n; i++) s += (double)in[i] * (double)((vg[i] + vfg[i]) - T(2) * m2 * gin[i]);
    return s / ((double)sigma * sigma * sigma);
}

// ------------------------------------------------------------------------------------
// ComputeNormal (computenormal.cu:17-43): per triangle n = normalize((v1 - v0) x (v2 - v0)),
// added to each corner's vertex with weight acos(e1 . e2) of that corner's two edges (mode 0,
// the raw edges as written, computenormal.cu:26-28) or of the unit edges (mode 1, the corner's
// interior angle); then each vertex sum is normalised.  The backward (a stub in the reference,
// computenormal.cu:56-65) is the exact derivative of that forward w.r.t. the positions.
// pinned: radial normals of a regular icosahedron and +z normals of a planar grid in both
// modes, backward by central differences (test_compute_normal_*).
template <class T>
static void cn_corner(const T* a, const T* b, const T* c, int mode, T& theta, T* dth_da, T* dth_db, T* dth_dc) {
    // corner at a with edges e1 = b - a, e2 = c - a
    T e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    T d = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    if (mode == 0) {
        theta = std::acos(d);
        if (dth_da) {
            T s = -T(1) / std::sqrt(T(1) - d * d);
            for (int i = 0; i < 3; i++) { dth_db[i] = s * e2[i]; dth_dc[i] = s * e1[i]; dth_da[i] = -(dth_db[i] + dth_dc[i]); }
        }
    } else {
        T l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]), l2 = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        T cs = d / (l1 * l2);
        theta = std::acos(cs);
        if (dth_da) {
            T s = -T(1) / std::sqrt(T(1) - cs * cs);
            for (int i = 0; i < 3; i++) {
                T dcs_db = e2[i] / (l1 * l2) - cs * e1[i] / (l1 * l1);
                T dcs_dc = e1[i] / (l1 * l2) - cs * e2[i] / (l2 * l2);
                dth_db[i] = s * dcs_db; dth_dc[i] = s * dcs_dc; dth_da[i] = -(dth_db[i] + dth_dc[i]);
            }
        }
    }
}

template <class T>
static void cn_face(const T* v0, const T* v1, const T* v2, T* n, T& len) {
    T e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]}, e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    T c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    T il = T(1) / len;
    n[0] = c[0] * il; n[1] = c[1] * il; n[2] = c[2] * il;
}

template <class T>
static void compute_normal_fwd(int V, const T* pos, int Tn, const u32* idx, int mode, T* out, T* sum) {
    std::vector<T> s((size_t)V * 3, T(0));
    for (int t = 0; t < Tn; t++) {
        const u32* f = idx + 3 * (size_t)t;
        const T* p[3] = {pos + 3 * (size_t)f[0], pos + 3 * (size_t)f[1], pos + 3 * (size_t)f[2]};
        T n[3], len;
        cn_face<T>(p[0], p[1], p[2], n, len);
        for (int c = 0; c < 3; c++) {
            T th;
            cn_corner<T>(p[c], p[(c + 1) % 3], p[(c + 2) % 3], mode, th, nullptr, nullptr, nullptr);
            for (int i = 0; i < 3; i++) s[3 * (size_t)f[c] + i] += n[i] * th;
        }
    }
    for (int v = 0; v < V; v++) {
        T* o = s.data() + 3 * (size_t)v;
        if (sum) for (int i = 0; i < 3; i++) sum[3 * (size_t)v + i] = o[i];
        T il = T(1) / std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        for (int i = 0; i < 3; i++) out[3 * (size_t)v + i] = o[i] * il;
    }
}

template <class T>
static void compute_normal_bwd(int V, const T* pos, int Tn, const u32* idx, int mode, const T* gout, T* gpos) {
    std::vector<T> s((size_t)V * 3), out((size_t)V * 3), gs((size_t)V * 3);
    compute_normal_fwd<T>(V, pos, Tn, idx, mode, out.data(), s.data());
    for (int v = 0; v < V; v++) {  // normalisation: dL/ds = (g - (g . o) o) / |s|
        const T* o = out.data() + 3 * (size_t)v;
        const T* g = gout + 3 * (size_t)v;
        T l = std::sqrt(s[3 * v] * s[3 * v] + s[3 * v + 1] * s[3 * v + 1] + s[3 * v + 2] * s[3 * v + 2]);
        T go = g[0] * o[0] + g[1] * o[1] + g[2] * o[2];
        for (int i = 0; i < 3; i++) gs[3 * (size_t)v + i] = (g[i] - go * o[i]) / l;
    }
    for (int t = 0; t < Tn; t++) {
        const u32* f = idx + 3 * (size_t)t;
        const T* p[3] = {pos + 3 * (size_t)f[0], pos + 3 * (size_t)f[1], pos + 3 * (size_t)f[2]};
        T n[3], len;
        cn_face<T>(p[0], p[1], p[2], n, len);
        T a[3] = {0, 0, 0};
        T gp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int c = 0; c < 3; c++) {
            T th, dA[3], dB[3], dC[3];
            cn_corner<T>(p[c], p[(c + 1) % 3], p[(c + 2) % 3], mode, th, dA, dB, dC);
            const T* G = gs.data() + 3 * (size_t)f[c];
            for (int i = 0; i < 3; i++) a[i] += th * G[i];
            T b = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];  // dL/dtheta of this corner
            for (int i = 0; i < 3; i++) {
                gp[c][i] += b * dA[i]; gp[(c + 1) % 3][i] += b * dB[i]; gp[(c + 2) % 3][i] += b * dC[i];
            }
        }
        // through n = cross / |cross|: h = (I - n n^T) a / |cross|; dL/dv0 = (v1 - v2) x h, ...
        T an = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
        T h[3] = {(a[0] - an * n[0]) / len, (a[1] - an * n[1]) / len, (a[2] - an * n[2]) / len};
        for (int c = 0; c < 3; c++) {
            const T* pb = p[(c + 1) % 3];
            const T* pc = p[(c + 2) % 3];
            T d[3] = {pb[0] - pc[0], pb[1] - pc[1], pb[2] - pc[2]};
            gp[c][0] += d[1] * h[2] - d[2] * h[1];
            gp[c][1] += d[2] * h[0] - d[0] * h[2];
            gp[c][2] += d[0] * h[1] - d[1] * h[0];
        }
        for (int c = 0; c < 3; c++)
            for (int i = 0; i < 3; i++) gpos[3 * (size_t)f[c] + i] += gp[c][i];
    }
}

// ------------------------------------------------------------------------------------
// Debug raster shadings (rasterize.cpp:292-512), from the rasterizer's output.
// Id hash (rasterize.cpp:430-451): the triangle id mixed with 0xbeefcade, 0xdeadcafe by
// Jenkins' mix, colour = (a, b, c) & 0x7fffff / 0x7fffff, alpha 1; background 0.
// Wireframe (rasterize.cpp:335-346): barycentrics var = (u, v, 1 - u - v) and
// d = fwidth(var / 2) = (|dvar/dX| + |dvar/dY|) / 2 from the analytic DB (GL takes finite
// differences over 2 x 2 quads), f = 1 / (1 + exp(16 (d - var))),
// e = 1 - min(f), colour (e, e, e, 1); background 0.
// parity unpinned against GL (fwidth's quad differences are the hardware's); the id hash's
// bits share the mix above.  pinned: wireframe edges and interiors (test_debug_raster).
extern "C" void orc_raster_idhash(size_t P, const float* rast, float* out) {
    for (size_t p = 0; p < P; p++) {
        float* o = out + 4 * p;
        int id = (int)rast[4 * p + 3] - 1;
        if (id < 0) { o[0] = o[1] = o[2] = o[3] = 0.f; continue; }
        u32 a = (u32)id, b = 0xbeefcadeu, c = 0xdeadcafeu;
        jenkins_mix(a, b, c);
        const u32 d = 0x007fffffu;
        o[0] = float(a & d) / float(d); o[1] = float(b & d) / float(d); o[2] = float(c & d) / float(d); o[3] = 1.f;
    }
}
extern "C" void orc_raster_wireframe(size_t P, const float* rast, const float* db, float* out) {
    for (size_t p = 0; p < P; p++) {
        float* o = out + 4 * p;
        const float* r = rast + 4 * p;
        const float* g = db + 4 * p;
        if ((int)r[3] - 1 < 0) { o[0] = o[1] = o[2] = o[3] = 0.f; continue; }
        float var[3] = {r[0], r[1], 1.f - r[0] - r[1]};
        float d[3] = {0.5f * (std::fabs(g[0]) + std::fabs(g[1])), 0.5f * (std::fabs(g[2]) + std::fabs(g[3])),
                      0.5f * (std::fabs(g[0] + g[2]) + std::fabs(g[1] + g[3]))};
        float fm = 1e30f;
        for (int i = 0; i < 3; i++) fm = std::min(fm, 1.f / (1.f + std::exp(16.f * (d[i] - var[i]))));
        float e = 1.f - fm;
        o[0] = o[1] = o[2] = e; o[3] = 1.f;
    }
}

// ------------------------------------------------------------------------------------
#define ORC_EXTRA(S, T)                                                                                                             \
extern "C" double orc_ssim_##S(int W, int H, int D, const T* pred, const T* target, T* grad, T* buffer, int mode) {                \
    return ssim<T>(W, H, D, pred, target, grad, buffer, mode); }                                                                   \
extern "C" void orc_filter_fwd_##S(int W, int H, int D, int C, T sigma, const T* in, T* out) { filter_fwd<T>(W, H, D, C, sigma, in, out); } \
extern "C" double orc_filter_bwd_##S(int W, int H, int D, int C, T sigma, const T* in, const T* gout, T* gin) {                    \
    return filter_bwd<T>(W, H, D, C, sigma, in, gout, gin); }                                                                      \
extern "C" void orc_compute_normal_fwd_##S(int V, const T* pos, int Tn, const u32* idx, int mode, T* out) {                         \
    compute_normal_fwd<T>(V, pos, Tn, idx, mode, out, nullptr); }                                                                  \
extern "C" void orc_compute_normal_bwd_##S(int V, const T* pos, int Tn, const u32* idx, int mode, const T* gout, T* gpos) {        \
    compute_normal_bwd<T>(V, pos, Tn, idx, mode, gout, gpos); }

ORC_EXTRA(f32, float)
ORC_EXTRA(f64, double)
