/* vcl.h -- C ABI of the MI355X differentiable-rendering library (libvcl_amd.so).
 *
 * Each op keeps the reference's POD parameter layout ("XKernelParams" / "XKernelGradParams",
 * passed by value to the kernels in the reference) with the same field names and meanings,
 * plus a stream and an int status.  Citations name the reference header each struct and
 * entry point replaces (236088/vcl, nvdiffrast/nvdiffrast/<file>:<line>).
 *
 * Conventions kept from the reference:
 *  - images are row-major with row 0 at the bottom (NDC y = -1); pixel index
 *    pidx = px + width * (py + height * pz), pz = layer ("depth");
 *  - a NULL grad pointer means "skip this gradient";
 *  - parameter grads accumulate (+=); grad.rast (interpolate), grad.uv (texture) and
 *    grad.in (antialias) are stored, not accumulated;
 *  - output grads are zeroed by the caller before the forward of a step (the reference's
 *    forward(XGradParams&) overloads); include/vcl_layers.hpp does this.
 * Additions: "depth" > 1 gives one clip-coordinate slice (projNum x 4 floats) per layer, so a
 * batch of views of one mesh renders in one call.  Scratch memory is passed in explicitly
 * (workspace pointers sized by the *_workspace_size queries), so no entry point allocates.
 * All entry points are asynchronous on the given stream and capturable in a HIP graph.
 */
#ifndef VCL_H_
#define VCL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* vcl_stream_t; /* hipStream_t */

enum {
    VCL_OK = 0,
    VCL_ERR_INVALID = 1,   /* bad dimensions or a required pointer is NULL */
    VCL_ERR_OOM = 2,       /* workspace too small / admission refused */
    VCL_ERR_HIP = 3,       /* a HIP runtime call failed */
    VCL_ERR_OVERFLOW = 4   /* device-side capacity exceeded (reported by *_status) */
};

#define VCL_TEX_MAX_MIP_LEVEL 16 /* TEX_MAX_MIP_LEVEL, buffer.h:4 */

const char* vcl_status_string(int status);
/* Library version and the device architecture the kernels were built for ("gfx950"). */
const char* vcl_build_info(void);

/* ---------------------------------------------------------------- Project
 * ProjectKernelParams / ProjectKernelGradParams (project.h:5-19); Project::forward/backward
 * (project.cu:69-84, 114-138).  out[v][i] = sum_j mat[4j+i] * [vec, 1]_j, mat column-major.
 * views > 1: mat holds views x 16 values and out views x vecNum x dimention. */
typedef struct vcl_project_params {
    int vecNum;
    int dimention;        /* 3 or 4 (homogeneous) */
    const float* vec;     /* vecNum x 3 */
    const float* mat;     /* views x 16 */
    float* out;           /* views x vecNum x dimention */
    int views;            /* >= 1 */
} vcl_project_params;

typedef struct vcl_project_grad_params {
    float* vec;           /* vecNum x 3, accumulated (summed over views), or NULL */
    float* mat;           /* views x 16, accumulated, or NULL (race-free here; project.cu:125-129) */
    const float* out;     /* views x vecNum x dimention */
} vcl_project_grad_params;

int vcl_project_forward(const vcl_project_params* p, vcl_stream_t stream);
int vcl_project_backward(const vcl_project_params* p, const vcl_project_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- Rasterize
 * RasterizeKernelParams (rasterize.h:6-18) plus projNum/idxNum of RasterizeParams
 * (rasterize.h:30-31); Rasterize::forward (rasterize.cpp:257-282, an OpenGL draw in the
 * reference, a tiled software rasterizer here) and Rasterize::backward (rasterize.cu:143-148).
 * out = (u, v, z/w, id+1) per pixel, outDB = (du/dX, du/dY, dv/dX, dv/dY) in pixel units.
 * Coverage: 8-bit sub-pixel fixed point, top-left fill rule, no culling; visibility: the
 * minimum of (z/w, triangle id).  Triangle ids must be < 2^24 (exact as float). */
typedef struct vcl_rasterize_params {
    int width;
    int height;
    int depth;            /* layers; layer z reads proj + z * projNum * 4 */
    int enableDB;
    float xs;             /* 2 / width  (rasterize.cpp:73) */
    float ys;             /* 2 / height (rasterize.cpp:74) */
    const float* proj;    /* depth x projNum x 4 clip coordinates */
    const uint32_t* idx;  /* idxNum x 3 */
    float* out;           /* depth x height x width x 4 */
    float* outDB;         /* depth x height x width x 4, or NULL when enableDB == 0 */
    int projNum;
    int idxNum;
    void* workspace;      /* device scratch of workspace_size bytes */
    size_t workspace_size;
} vcl_rasterize_params;

typedef struct vcl_rasterize_grad_params {
    float* proj;          /* depth x projNum x 4, accumulated (x, y, w) */
    const float* out;     /* dL/d(u,v): float2 per pixel, packed (rasterize.cu:94) */
    const float* outDB;   /* ignored, as in the reference (rasterize.cu:96-98) */
} vcl_rasterize_grad_params;

size_t vcl_rasterize_workspace_size(int width, int height, int depth, int projNum, int idxNum);
int vcl_rasterize_forward(const vcl_rasterize_params* p, vcl_stream_t stream);
int vcl_rasterize_backward(const vcl_rasterize_params* p, const vcl_rasterize_grad_params* g, vcl_stream_t stream);
/* Device-side status of the last forward on this workspace (VCL_OK or VCL_ERR_OVERFLOW);
 * synchronises the stream. */
int vcl_rasterize_status(const vcl_rasterize_params* p, vcl_stream_t stream);
/* Testing hook (no counterpart in the reference): fix the tile-bin capacity, in entries, of
 * workspaces sized after this call (0 restores the default sizing).  A forward whose bins
 * would pass the capacity takes the slower scan-every-piece path for the affected tiles and
 * stays exact. */
void vcl_rasterize_set_bin_capacity(size_t entries);

/* ---------------------------------------------------------------- Interpolate
 * InterpolateKernelParams / InterpolateKernelGradParams (interpolate.h:7-29);
 * Interpolate::forward/backward (interpolate.cu:51-60, 145-150). */
typedef struct vcl_interpolate_params {
    int width;
    int height;
    int depth;
    int dimention;
    int enableDA;
    const float* rast;
    const float* rastDB;
    const float* attr;    /* attrNum x dimention */
    const uint32_t* idx;  /* triangles x 3 */
    float* out;           /* P x dimention (background written 0) */
    float* outDA;         /* P x dimention x 2 */
    int attrNum;          /* vertex count of attr */
} vcl_interpolate_params;

typedef struct vcl_interpolate_grad_params {
    const float* out;
    const float* outDA;   /* ignored, as in the reference (interpolate.cu:114-143) */
    float* rast;          /* float2 per pixel, stored for covered pixels */
    float* rastDB;        /* ignored */
    float* attr;          /* accumulated, or NULL */
} vcl_interpolate_grad_params;

int vcl_interpolate_forward(const vcl_interpolate_params* p, vcl_stream_t stream);
int vcl_interpolate_backward(const vcl_interpolate_params* p, const vcl_interpolate_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- Texturemap
 * TexturemapKernelParams / TexturemapKernelGradParams (texturemap.h:8-31);
 * Texturemap::forward/backward (texturemap.cu:116-123, 321-327).  Trilinear mip sampling when
 * miplevel > 1, bilinear level 0 otherwise.  Levels are HWC, level l is
 * (texwidth >> l) x (texheight >> l) x channel. */
typedef struct vcl_texturemap_params {
    int width;
    int height;
    int depth;
    int texwidth;
    int texheight;
    int channel;
    int miplevel;
    const float* rast;
    const float* uv;      /* P x 2 */
    const float* uvDA;    /* P x 4 */
    const float* texture[VCL_TEX_MAX_MIP_LEVEL];
    float* out;           /* P x channel (background written 0) */
} vcl_texturemap_params;

enum { VCL_TEX_GRAD_DA_REF = 0, VCL_TEX_GRAD_DA_EXACT = 1 };

typedef struct vcl_texturemap_grad_params {
    const float* out;
    float* uv;            /* P x 2, stored, or NULL */
    float* uvDA;          /* P x 4, stored, or NULL */
    float* texture[VCL_TEX_MAX_MIP_LEVEL]; /* accumulated, or NULL */
    int daMode;           /* VCL_TEX_GRAD_DA_REF: formula of texturemap.cu:160-197,282;
                             VCL_TEX_GRAD_DA_EXACT: exact d level / d uvDA */
} vcl_texturemap_grad_params;

int vcl_texturemap_forward(const vcl_texturemap_params* p, vcl_stream_t stream);
int vcl_texturemap_backward(const vcl_texturemap_params* p, const vcl_texturemap_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- Texture (mip utilities)
 * Texture / TextureGrad (buffer.h:54-78): buildMIP (buffer.cu:474-484), clamp (:588-594),
 * TextureGrad::clear (:607-613), gradSumup (:638-648). */
typedef struct vcl_texture {
    int width;
    int height;
    int channel;
    int miplevel;
    float* texture[VCL_TEX_MAX_MIP_LEVEL];
} vcl_texture;

/* Mip level count of Texture::init (buffer.cu:380-385). */
int vcl_texture_mip_levels(int width, int height, int requested);
int vcl_texture_build_mip(const vcl_texture* t, vcl_stream_t stream);
int vcl_texture_clamp(const vcl_texture* t, float min, float max, vcl_stream_t stream);
/* grad holds the per-level gradients.  quarter = 1 scales each folded level by 1/4 (exact
 * adjoint of buildMIP); quarter = 0 adds it unscaled as the reference does. */
int vcl_texture_grad_sumup(const vcl_texture* grad, int quarter, vcl_stream_t stream);
int vcl_texture_grad_clear(const vcl_texture* grad, vcl_stream_t stream);

/* ---------------------------------------------------------------- Optimizer / Adam
 * OptimizeKernelParams (optimize.h:5-13), AdamHyperParams (optimize.h:50-59), adamStep
 * (optimize.cu:95-107).  The caller advances rhomt/rhovt before each step
 * (Adam::step, optimize.cu:109-115); the update is computed in double. */
typedef struct vcl_optimize_params {
    float* param;
    const float* grad;
    int size;
    int width;
    int height;
    int depth;
} vcl_optimize_params;

typedef struct vcl_adam_hyper_params {
    double rhom;
    double rhov;
    double rhomt;
    double rhovt;
    double eta;
    double eps;
    float* m;
    float* v;
} vcl_adam_hyper_params;

int vcl_adam_step(const vcl_optimize_params* p, const vcl_adam_hyper_params* h, vcl_stream_t stream);
/* Multi-tensor Adam: all n tensors in one launch. */
int vcl_adam_step_multi(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, vcl_stream_t stream);
/* Fused texture update: gradSumup -> Adam(level 0) -> TextureGrad::clear -> clamp -> buildMIP
 * (the order of preset_primitives.cpp:200-202 with buffer.cu:588-648). */
int vcl_texture_adam_update(const vcl_texture* tex, const vcl_texture* grad, const vcl_adam_hyper_params* h,
                            float clamp_min, float clamp_max, int quarter, vcl_stream_t stream);

/* ---------------------------------------------------------------- Antialias
 * AntialiasKernelParams / AntialiasKernelGradParams (antialias.h:7-27);
 * Antialias::forward/backward (antialias.cu:97-103, 302-313).  Race-free: out = in plus all
 * edge-leak terms (the reference's copy races its atomics); the backward uses the forward's
 * pixel selection and the exact derivative including w. */
typedef struct vcl_antialias_params {
    int width;
    int height;
    int depth;
    int channel;
    float s[2];           /* (width/2, height/2) */
    const float* proj;    /* depth x projNum x 4 */
    const uint32_t* idx;
    const float* rast;
    const float* in;
    float* out;
    int projNum;
} vcl_antialias_params;

typedef struct vcl_antialias_grad_params {
    float* proj;          /* accumulated, or NULL */
    float* in;            /* stored */
    const float* out;
} vcl_antialias_grad_params;

int vcl_antialias_forward(const vcl_antialias_params* p, vcl_stream_t stream);
int vcl_antialias_backward(const vcl_antialias_params* p, const vcl_antialias_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- Material
 * NormalAxisKernelParams (material.h:9-25), NormalAxis::forward (material.cu:71-77).
 * out = rot3x3 * normalize(barycentric normal) -- the effective reference semantics
 * (its normal-map branch is inverted, material.cu:40-48); normalmap is ignored. */
typedef struct vcl_normal_axis_params {
    int width;
    int height;
    int depth;
    const float* rast;
    const float* rot;         /* 16 values, column-major (device) */
    const float* normal;      /* N x 3 */
    const uint32_t* normalidx;
    const float* pos;
    const uint32_t* posidx;
    const float* texel;
    const uint32_t* texelidx;
    const float* normalmap;
    float* out;               /* P x 3 (background written 0) */
} vcl_normal_axis_params;

int vcl_normal_axis_forward(const vcl_normal_axis_params* p, vcl_stream_t stream);

/* ViewAxisKernelParams (material.h:52-65), ViewAxis::forward (material.cu:110-118).
 * pvinv = inverse(mat3(projection) * mat3(view)) is formed on the host from the host
 * matrices, as in the reference; one pair of matrices per layer. */
typedef struct vcl_view_axis_params {
    int width;
    int height;
    int depth;
    const float* rast;        /* unused, kept for layout */
    const float* rot;         /* unused, kept for layout */
    const float* view;        /* host, depth x 16 column-major */
    const float* projection;  /* host, depth x 16 column-major */
    float* pvinv;             /* unused (the inverse travels as a kernel argument) */
    float* out;               /* P x 3 */
} vcl_view_axis_params;

int vcl_view_axis_forward(const vcl_view_axis_params* p, vcl_stream_t stream);

/* SGSpecularKernelParams / SGSpecularGradKernelParams (material.h:94-122);
 * SGSpecular::forward/backward (material.cu:185-194, 251-256). */
typedef struct vcl_sg_specular_params {
    int width;
    int height;
    int depth;
    const float* rast;
    const float* normal;
    const float* view;
    const float* roughness;   /* P */
    float ior;
    float* outAxis;           /* P x 3 */
    float* outSharpness;      /* P */
    float* outAmplitude;      /* P */
} vcl_sg_specular_params;

typedef struct vcl_sg_specular_grad_params {
    float* normal;            /* unused, as in the reference */
    float* view;              /* unused */
    float* roughness;         /* accumulated */
    float ior;
    const float* outAxis;     /* unused, as in the reference */
    const float* outSharpness;
    const float* outAmplitude;
} vcl_sg_specular_grad_params;

int vcl_sg_specular_forward(const vcl_sg_specular_params* p, vcl_stream_t stream);
int vcl_sg_specular_backward(const vcl_sg_specular_params* p, const vcl_sg_specular_grad_params* g, vcl_stream_t stream);

/* SphericalGaussianKernelParams / ...GradKernelParams (material.h:144-183);
 * SphericalGaussian::forward/backward (material.cu:325-334, 440-445).  Lobe gradients are
 * reduced in a fixed order (the reference accumulates them non-atomically), so results are
 * deterministic; the workspace holds the per-block partials. */
typedef struct vcl_spherical_gaussian_params {
    int width;
    int height;
    int depth;
    int channel;
    const float* rast;
    int sgnum;
    const float* axis;        /* sgnum x 3 */
    const float* sharpness;   /* sgnum */
    const float* amplitude;   /* sgnum x channel */
    const float* normal;
    const float* view;        /* unused, kept for layout */
    const float* diffuse;     /* P x channel */
    const float* specAxis;
    const float* specSharpness;
    const float* specAmplitude;
    float* out;
    float* outDiffenv;
    float* outSpecenv;
} vcl_spherical_gaussian_params;

typedef struct vcl_spherical_gaussian_grad_params {
    float* normal;            /* P x 3, accumulated, or NULL */
    float* diffuse;           /* P x channel, accumulated, or NULL */
    float* axis;              /* sgnum x 3, accumulated, or NULL */
    float* sharpness;         /* sgnum, accumulated, or NULL */
    float* amplitude;         /* sgnum x channel, accumulated, or NULL */
    float* specAxis;          /* P x 3, accumulated, or NULL */
    float* specSharpness;     /* P, accumulated, or NULL */
    float* specAmplitude;     /* P, accumulated, or NULL */
    const float* out;
    void* workspace;          /* >= vcl_spherical_gaussian_workspace_size() bytes */
    size_t workspace_size;
} vcl_spherical_gaussian_grad_params;

size_t vcl_spherical_gaussian_workspace_size(int width, int height, int depth, int channel, int sgnum);
int vcl_spherical_gaussian_forward(const vcl_spherical_gaussian_params* p, vcl_stream_t stream);
int vcl_spherical_gaussian_backward(const vcl_spherical_gaussian_params* p, const vcl_spherical_gaussian_grad_params* g,
                                    vcl_stream_t stream);

/* ---------------------------------------------------------------- Loss
 * LossKernelParams (loss.h:5-15), MSELoss::backward (loss.cu:41-57): grad = predict - target,
 * loss = 0.5 * sum(grad^2) written to a device float (no host sync); fixed-order reduction. */
typedef struct vcl_loss_params {
    const float* predict;
    const float* target;
    float* grad;
    float* buffer;            /* device scratch of vcl_loss_workspace_size() bytes */
    int width;
    int height;
    int depth;
    int size;                 /* element count */
    float* loss;              /* device, 1 float */
} vcl_loss_params;

size_t vcl_loss_workspace_size(int size);
int vcl_mse_loss_backward(const vcl_loss_params* p, vcl_stream_t stream);
/* The same MSE loss without allocation or workspace argument (no counterpart in the
 * reference): for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  p->buffer is ignored; the block partials live in
 * a 1 MiB device array of the library, one per GPU, held until the process ends, so sizes up
 * to vcl_mse_loss_noalloc_max_size() elements (2^28) are accepted and larger ones return
 * VCL_ERR_OOM.  Results equal vcl_mse_loss_backward's.  Calls that overlap on one GPU share
 * that array and race. */
size_t vcl_mse_loss_noalloc_max_size(void);
int vcl_mse_loss_backward_noalloc(const vcl_loss_params* p, vcl_stream_t stream);

/* ---------------------------------------------------------------- Further losses
 * SSIMLoss::backward (loss.h:18-20, loss.cu:93-174) on vcl_loss_params: per element (HWC with
 * `depth` values per pixel, pidx = (px + width py) depth + pz) the SSIM of its (2k+1)^2 window,
 * k = 5, clipped to the image, means and (co)variances over k^2 = 25 (loss.cu:155),
 * c1 = 1e-4, c2 = 1e-3.  buffer: device scratch of vcl_ssim_workspace_size bytes whose first
 * width*height*depth floats hold 1 - SSIM on return (the reference's loss.buffer);
 * *loss = sum(1 - SSIM) / (4 size) (the reference divides by the byte size, loss.cu:173).
 * gradMode VCL_SSIM_GRAD_REF: loss.cu:146-149 as written (the centre element's term of its own
 * window, with 1 - SSIM where the derivative has SSIM); VCL_SSIM_GRAD_EXACT: the derivative of
 * sum(1 - SSIM) over every window that holds the element.  grad is stored. */
enum { VCL_SSIM_GRAD_REF = 0, VCL_SSIM_GRAD_EXACT = 1 };
size_t vcl_ssim_workspace_size(int width, int height, int depth);
int vcl_ssim_loss_backward(const vcl_loss_params* p, int gradMode, vcl_stream_t stream);
/* MSELoss::textureloss (loss.cu:59-91): squared differences outside the masked block
 * ((px < W/4 or W/2 <= px) and (py < W/4 or W/2 <= py) are zero), *loss = sum / (size / 2);
 * no gradient.  buffer: vcl_loss_workspace_size(size) bytes plus size floats. */
size_t vcl_texture_loss_workspace_size(int size);
int vcl_mse_texture_loss(const vcl_loss_params* p, vcl_stream_t stream);

/* ---------------------------------------------------------------- Further optimizers
 * SGD (optimize.h:36-48, optimize.cu:64-82): param += -eta * grad (AddNaNcheck).
 * Nadam (optimize.h:71-98, optimize.cu:118-167): the kernel of Nadam::step with the host state
 * as given (the caller advances it as Nadam::step does, optimize.cu:157-163).
 * Optimizer::randomParams / clampParams (optimize.cu:30-60); the reference draws the seed with
 * rand(), here it is an argument. */
typedef struct vcl_sgd_hyper_params {
    double eta;
} vcl_sgd_hyper_params;
typedef struct vcl_nadam_hyper_params {
    double mupow, mupowt, mu, mu0, mu1, rho, alpha, mut0, mut1, rhot, eps;
    float* m;
    float* v;
} vcl_nadam_hyper_params;
int vcl_sgd_step(const vcl_optimize_params* p, const vcl_sgd_hyper_params* h, vcl_stream_t stream);
int vcl_nadam_step(const vcl_optimize_params* p, const vcl_nadam_hyper_params* h, vcl_stream_t stream);
int vcl_optimize_random_params(const vcl_optimize_params* p, uint32_t seed, float min, float max, vcl_stream_t stream);
int vcl_optimize_clamp_params(const vcl_optimize_params* p, float min, float max, vcl_stream_t stream);

/* ---------------------------------------------------------------- Buffer utilities
 * The kernels behind Buffer / Attribute / Texture / SGBuffer helpers (buffer.cu:3-110,
 * :313-361, :410-447, :543-594, :669-811).  n counts floats; dst may equal src. */
int vcl_buffer_linear(float* dst, const float* src, size_t n, float w, float b, vcl_stream_t stream);        /* linerKernel */
int vcl_buffer_add_random(float* dst, size_t n, float min, float max, uint32_t seed, vcl_stream_t stream);    /* randomKernel */
int vcl_buffer_clamp(float* dst, const float* src, size_t n, float min, float max, vcl_stream_t stream);    /* clampKernel */
int vcl_buffer_step(float* dst, const float* src, size_t n, float threshold, vcl_stream_t stream);          /* stepKernel */
/* normalizeKernel: each of `rows` rows of `dimention` floats times 1/|src row| (zero rows kept) */
int vcl_buffer_normalize(float* dst, const float* src, int rows, int dimention, vcl_stream_t stream);
/* setValueKernel: every row of `dimention` (<= 16) floats set to value (host array) */
int vcl_buffer_set_value(float* dst, int rows, int dimention, const float* value, vcl_stream_t stream);
/* sphericalRandomKernel: count unit directions (3 floats each) */
int vcl_buffer_spherical_random(float* dst, int count, uint32_t seed, vcl_stream_t stream);
/* Attribute::distanceError (buffer.cu:321-344): mean |predict - target| over num rows of
 * `dimention`, into *out (device float), in a fixed order. */
int vcl_attribute_distance_error(const float* predict, const float* target, int num, int dimention, float* out, vcl_stream_t stream);
/* Texture::bilinearDownsampling (buffer.cu:410-447): levels 1.. from level 0 by the [1 3 3 1]/8
 * separable filter, channels 0..min(channel, 3)-1 (the reference writes three). */
int vcl_texture_bilinear_downsampling(const vcl_texture* t, vcl_stream_t stream);
/* SGBuffer::bake (buffer.cu:734-761): level 0 of t = sum of the lobes over equirectangular
 * directions (t->channel = the lobes' channel count); the other levels are cleared first, as
 * Texture::liner(texture, 0, 0) does. */
int vcl_sg_bake(const float* axis, const float* sharpness, const float* amplitude, int num, int channel, const vcl_texture* t,
                vcl_stream_t stream);
/* SGBuffer::normalize (buffer.cu:685-704): unit axes, sharpness in [0, 8], amplitude in [0, 1]. */
int vcl_sg_normalize(float* axis, float* sharpness, float* amplitude, int num, int channel, vcl_stream_t stream);
/* SGBufferGrad::disperse (buffer.cu:771-804): axes pushed apart; buffer holds num x 3 floats.
 * Amplitude differences are taken from the amplitudes (the reference indexes sharpness). */
int vcl_sg_disperse(float* axis, const float* sharpness, const float* amplitude, int num, int channel, float* buffer,
                    vcl_stream_t stream);

/* ---------------------------------------------------------------- Noise
 * NoiseKernelParams (extention.h:5-14), Noise::forward (extention.cu:13-30):
 * out = in + intensity (U(pidx, seed + i, 0xcafef00d) - in); the seed is an argument (the
 * reference draws it with rand()). */
typedef struct vcl_noise_params {
    int width, height, depth, channel;
    float intensity;
    const float* in;
    float* out;
} vcl_noise_params;
int vcl_noise_forward(const vcl_noise_params* p, uint32_t seed, vcl_stream_t stream);

/* ---------------------------------------------------------------- Filter
 * FilterKernelParams / FilterKernelGradParams (filter.h:5-31), Filter::forward/backward
 * (filter.cu:62-186): separable Gaussian blur, horizontal then vertical, taps outside the image
 * dropped, k = min(ceil(4 sigma), 127).  sigma stays on the device and the weights are formed
 * there (the reference copies sigma to the host each call), so no call synchronises.  The
 * backward stores grad.in (the same blur of grad.out) and *grad.sigma = d loss / d sigma,
 * computed exactly (the reference's FilterBackwardKernelFinal adds each pixel's term once per
 * vertical tap and uses 2 / sigma for 2 m2 / sigma^3, filter.cu:141-157,182).
 * filter: 256 floats of scratch; buf, bufx, bufy: width*height*depth*channel floats each. */
typedef struct vcl_filter_params {
    int width, height, depth, channel;
    int k;                  /* unused: derived from sigma on the device */
    const float* sigma;     /* device, 1 float */
    float* filter;          /* device scratch, 256 floats */
    const float* in;
    float* buf;             /* device scratch */
    float* out;
} vcl_filter_params;
typedef struct vcl_filter_grad_params {
    float* in;              /* stored */
    float* sigma;           /* device, 1 float, stored */
    float* filter;          /* device scratch, 256 floats */
    float* buf;
    float* bufx;
    float* bufy;
    const float* out;
} vcl_filter_grad_params;
int vcl_filter_forward(const vcl_filter_params* p, vcl_stream_t stream);
int vcl_filter_backward(const vcl_filter_params* p, const vcl_filter_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- ComputeNormal
 * ComputeNormalKernelParams (computenormal.h:5-19), ComputeNormal::forward (computenormal.cu:
 * 17-43): per triangle the unit face normal added to its vertices with the weight acos(e1.e2)
 * of each corner's edges -- the raw edges as written (VCL_NORMAL_ANGLE_REF,
 * computenormal.cu:26-28) or the unit edges, i.e. the interior angle (VCL_NORMAL_ANGLE_EXACT) --
 * then normalised.  The backward (a stub in the reference, computenormal.cu:56-65) is the exact
 * derivative w.r.t. the positions.  sum: posNum x 3 floats of device scratch that keep the
 * unnormalised sums for the backward. */
enum { VCL_NORMAL_ANGLE_REF = 0, VCL_NORMAL_ANGLE_EXACT = 1 };
typedef struct vcl_compute_normal_params {
    int posNum;
    const float* pos;
    const uint32_t* vao;
    int vaoNum;
    const float* mat;       /* unused, kept for layout */
    float* out;             /* posNum x 3 */
    float* sum;             /* posNum x 3 device scratch */
    int angleMode;
} vcl_compute_normal_params;
typedef struct vcl_compute_normal_grad_params {
    float* pos;             /* posNum x 3, accumulated */
    const float* out;       /* posNum x 3 */
} vcl_compute_normal_grad_params;
int vcl_compute_normal_forward(const vcl_compute_normal_params* p, vcl_stream_t stream);
int vcl_compute_normal_backward(const vcl_compute_normal_params* p, const vcl_compute_normal_grad_params* g, vcl_stream_t stream);

/* ---------------------------------------------------------------- Debug raster shadings
 * Rasterize::wireframeinit / idhashinit + drawforward (rasterize.cpp:292-512), from a forward's
 * output (rast, and DB for the wireframe) into out (P x 4):
 * wireframe: e = 1 - min_i 1 / (1 + exp(16 (d_i - b_i))) over the barycentrics b = (u, v, 1-u-v)
 * with d = (|db/dX| + |db/dY|) / 2 from the analytic DB (GL's fwidth differences 2 x 2 quads),
 * colour (e, e, e, 1); idhash: the triangle id mixed with 0xbeefcade, 0xdeadcafe by the
 * reference's integer mix, colour (a, b, c) & 0x7fffff / 0x7fffff, alpha 1.  Background 0. */
int vcl_rasterize_wireframe(int width, int height, int depth, const float* rast, const float* rastDB, float* out, vcl_stream_t stream);
int vcl_rasterize_idhash(int width, int height, int depth, const float* rast, float* out, vcl_stream_t stream);

/* ---------------------------------------------------------------- On-disk formats (host)
 * Host-only readers and writers, usable without a GPU; returned arrays are malloc'd and freed
 * by vcl_obj_free / vcl_image_free.  include/vcl_layers.hpp uploads them for
 * Attribute::loadOBJ, Texture::loadBMP and SGBuffer::loadTXT. */
typedef struct vcl_obj_mesh {
    int posNum, texelNum, normalNum, indexNum;
    float* pos;              /* posNum x 3 */
    float* texel;            /* texelNum x 2, or NULL */
    float* normal;           /* normalNum x 3, or NULL */
    uint32_t* posIndex;      /* indexNum x 3 (0-based) */
    uint32_t* texelIndex;    /* indexNum x 3, or NULL */
    uint32_t* normalIndex;   /* indexNum x 3, or NULL */
} vcl_obj_mesh;
/* Attribute::loadOBJ (buffer.cu:203-306): token-driven v / vt / vn / f records, faces in the
 * form fixed by the records read so far (v/vt/vn, v/vt, v//vn or v); VCL_ERR_INVALID when the
 * file cannot be opened or a record does not parse (the reference prints and returns). */
int vcl_load_obj(const char* path, vcl_obj_mesh* out);
void vcl_obj_free(vcl_obj_mesh* m);
typedef struct vcl_image {
    int width, height, channel;
    float* data;             /* width x height x channel, row 0 at the bottom */
} vcl_image;
/* Texture::loadBMP's decode (buffer.cu:486-541): header fields at 0x0A, 0x12, 0x16, 0x1C, 0x22;
 * pixel bytes read contiguously (no row padding), per pixel in reversed order (BGR -> RGB),
 * divided by 255. */
int vcl_load_bmp(const char* path, vcl_image* out);
void vcl_image_free(vcl_image* im);
/* The exact inverse of vcl_load_bmp's decode (values clamped to [0, 1], times 255, rounded);
 * channel 1, 3 or 4. */
int vcl_save_bmp(const char* path, const float* data, int width, int height, int channel);
/* Portable float map ("Pf" for 1 channel, "PF" otherwise), rows bottom first. */
int vcl_save_pfm(const char* path, const float* data, int width, int height, int channel);
/* SGBuffer::loadTXT (buffer.cu:706-732): num lines of axis (3), sharpness, amplitude (3). */
int vcl_load_sg_txt(const char* path, int num, float* axis, float* sharpness, float* amplitude);

/* ---------------------------------------------------------------- Rotation / Camera (host)
 * Host math of Rotation and Camera (matrix.h:1-96, matrix.cu:3-239), usable without a GPU.
 * Quaternions are 4 floats in glm's storage order (x, y, z, w); matrices are 16 floats
 * column-major (glm::mat4).  include/vcl_matrix.hpp holds the same math header-only for C++
 * callers; include/vcl_layers.hpp builds the Rotation / Camera classes on it. */
/* getRotation (matrix.cu:43-53) */
int vcl_rotation_matrix(const float* q, float* out16);
/* getView (matrix.cu:175-180) */
int vcl_view_matrix(const float* q, const float* eye, float* out16);
/* Camera::setCam's look-at quaternion (matrix.cu:182-192) */
int vcl_look_at_quat(const float* eye, const float* center, const float* up, float* q_out);
/* Camera::forward's projection (matrix.cu:211-219) */
int vcl_projection_matrix(float size, float aspect, float znear, float zfar, float* out16);
/* Camera::forward (matrix.cu:209-222): view, projection and mat = projection * view * rotation. */
int vcl_camera_matrices(const float* q, const float* eye, float size, float aspect, float znear, float zfar,
                        const float* rotation16, float* view16, float* projection16, float* mat16);
/* Camera::backward (matrix.cu:234-239).  exact = 0: the reference's products
 * (G V^T P^T, R^T G P^T, R^T V^T G); exact = 1: the derivatives of mat = P V R. */
int vcl_camera_backward_host(const float* gmat16, const float* rotation16, const float* view16, const float* projection16,
                             int exact, float* grot16, float* gview16, float* gproj16);
/* Rotation::backward (matrix.cu:85-99).  exact = 0: the reference's sums (half the derivative);
 * exact = 1: the derivative of getRotation w.r.t. q. */
int vcl_rotation_backward_host(const float* q, const float* grot16, int exact, float* gq);
/* Rotation::init(..., eta, rhom, rhov, eps) state and Rotation::step (matrix.cu:101-137):
 * Adam on the quaternion in double with the reference's rhom^(t+1) bias correction, then
 * renormalisation.  q is updated in place. */
typedef struct vcl_rotation_adam {
    int it;
    double rhom, rhov, rhomt, rhovt, eta, eps;
    float m[4];
    float v[4];
} vcl_rotation_adam;
int vcl_rotation_adam_step(vcl_rotation_adam* a, float* q, const float* gq);

/* ---------------------------------------------------------------- Fit engine
 * One inverse-rendering request (the chain of preset_primitives.cpp:160-207 generalised to
 * vertex positions, mip texture and SG lobes): project -> rasterize(+DB) -> interpolate(uv)
 * -> NormalAxis -> ViewAxis -> texture -> SGSpecular -> SphericalGaussian -> antialias ->
 * MSE, backward, Adam.  All state stays resident on the device between steps. */
typedef struct vcl_fit_desc {
    int width, height, views;         /* views share mesh, texture and lobes */
    int V, T;                         /* vertices, triangles */
    int texwidth, texheight, channel, miplevel;
    int sgnum;
    const float* pos;                 /* host V x 3 (initial) */
    const float* normals;             /* host V x 3 */
    const float* uv;                  /* host V x 2 */
    const uint32_t* idx;              /* host T x 3 */
    const float* texture;             /* host, all levels back to back */
    const float* axis;                /* host K x 3 */
    const float* sharpness;           /* host K */
    const float* amplitude;           /* host K x channel */
    const float* mats;                /* host views x 16 (P*V*R) */
    const float* pvinv;               /* host views x 9 */
    const float* rot;                 /* host 16 */
    const float* target;              /* host views x H x W x channel, or NULL */
    float roughness, ior;
    double lr, beta1, beta2, eps;
    int fold_quarter;                 /* 1: exact mip-grad fold (default) */
    int op_by_op;                     /* 0: fused passes (interpolate + texture, NormalAxis +
                                         ViewAxis + SGSpecular, texture + interpolate backward);
                                         1: the reference's op-by-op sequence of separate launches.
                                         Both give the same results. */
} vcl_fit_desc;

typedef struct vcl_fit vcl_fit;       /* opaque */

int vcl_fit_create(const vcl_fit_desc* d, vcl_fit** out);
/* Device bytes of one request's arena (every buffer of vcl_fit_create_in); 0 if d is invalid. */
size_t vcl_fit_arena_bytes(const vcl_fit_desc* d);
/* vcl_fit_create inside caller-provided device memory (at least vcl_fit_arena_bytes(d) bytes;
 * it stays the caller's: vcl_fit_destroy does not free it).  Zeroes the arena and uploads the
 * host arrays on `stream`, then synchronises it.  No counterpart in the reference (whose init
 * calls cudaMalloc per buffer and never frees); it is what a serving pool needs. */
int vcl_fit_create_in(const vcl_fit_desc* d, void* arena, size_t arena_bytes, vcl_stream_t stream, vcl_fit** out);
int vcl_fit_destroy(vcl_fit* f);
/* Render the current parameters and store the image as the fit target. */
int vcl_fit_set_target_from_render(vcl_fit* f, vcl_stream_t stream);
/* One fwd+bwd+Adam iteration on the stream. */
int vcl_fit_step(vcl_fit* f, vcl_stream_t stream);
/* The same iteration replayed from a HIP graph (captured on the first call, on a stream of
 * the fit's own; ordered after prior work on `stream`, and later work on `stream` waits for
 * it).  Equivalent to vcl_fit_step; replaces main.cpp's per-op launch loop (main.cpp:39-49)
 * with one launch per iteration.  No counterpart in the reference API. */
int vcl_fit_step_graph(vcl_fit* f, vcl_stream_t stream);
/* Device pointers of named buffers ("pos", "texture", "axis", "sharpness", "amplitude",
 * "image", "loss", "clip", "rast", "db", "grad_pos", ...); NULL if unknown. */
void* vcl_fit_buffer(vcl_fit* f, const char* name);
size_t vcl_fit_device_bytes(const vcl_fit* f);
/* Copy bytes between a named device buffer and host memory; synchronises the stream. */
int vcl_fit_read(vcl_fit* f, const char* name, void* host, size_t bytes, vcl_stream_t stream);
int vcl_fit_write(vcl_fit* f, const char* name, const void* host, size_t bytes, vcl_stream_t stream);
/* Per-op HIP-event timing of the following steps (up to max_launches launches per op; only the
 * op named `only` when it is not NULL), the per-op timing the reference prints only as
 * whole-step fps (main.cpp:39-49).  _end syncs and returns, per op, its name
 * (names[i*name_len ...]), summed milliseconds and launch count. */
int vcl_fit_profile_begin(vcl_fit* f, int max_launches, const char* only);
int vcl_fit_profile_end(vcl_fit* f, int max_ops, int* nops, char* names, int name_len, double* total_ms, int* count);

/* ---------------------------------------------------------------- Tuning
 * Kernel schedule parameters (no counterpart in the reference): every value gives the same
 * results (the raster is exact for any tiling; texel gradients differ only in float summation
 * order), only the speed changes.  vcl_set_tuning applies to launches that follow it in this
 * process; vcl_fit_autotune picks them for one request. */
typedef struct vcl_tuning {
    int raster_tile;   /* fine-raster tile side in pixels: 8, 16 (default) or 32 */
    int bin_chunk;     /* pieces per fine-raster work item ("bin depth"): 64, 128 or 256 (default) */
    int splat_rows;    /* texture backward: rows of the 16-pixel-wide region whose texel-gradient
                          splats share one LDS table ("texel-splat width"): 16 (default), 32 or 64 */
} vcl_tuning;

int vcl_set_tuning(const vcl_tuning* t);
int vcl_get_tuning(vcl_tuning* t);
/* Tuning of one request's steps (NULL: follow the process-wide tuning); drops its graph. */
int vcl_fit_set_tuning(vcl_fit* f, const vcl_tuning* t);
int vcl_fit_get_tuning(const vcl_fit* f, vcl_tuning* t);
/* Autotuner: times `steps` steps of candidate tunings on this request (the current tuning,
 * then raster tile x bin chunk, then texel-splat rows; every candidate from the same snapshot
 * of the request's parameters and Adam state, which is restored afterwards).  A candidate
 * replaces the incumbent only when at least 3 % faster.  Keeps the winner on the request and
 * returns it with its ms per step.  Allocates a temporary snapshot of the persistent state and
 * synchronises the device first. */
int vcl_fit_autotune(vcl_fit* f, int steps, vcl_tuning* best, double* best_ms);
/* The same without allocation or device-wide synchronisation: the snapshot lives in caller
 * memory of at least vcl_fit_autotune_scratch_bytes(f) bytes, the trials run on `stream` (after
 * the work already queued on it), and only that stream is synchronised at the end.  The
 * server uses it with memory from its budgeted pool. */
size_t vcl_fit_autotune_scratch_bytes(const vcl_fit* f);
int vcl_fit_autotune_in(vcl_fit* f, int steps, void* scratch, size_t scratch_bytes, vcl_stream_t stream, vcl_tuning* best, double* best_ms);

/* ---------------------------------------------------------------- Groups
 * One step of several requests in shared launches (no counterpart in the reference, whose only
 * batch dimension is the layer count of one draw: one mesh, one texture, one resolution,
 * rasterize.h:42).  Requests of any size (mesh, resolution, texture, lobe count) whose steps
 * launch the same kernels with the same workgroup shape (equal vcl_fit_step_signature) form a
 * group.  A group step is one launch per kernel of the step for all its requests: each launch
 * reads a device-side work table (every request's workgroup range, grid and arguments), and
 * each workgroup runs the kernel for its own request.  Results equal each request's own
 * vcl_fit_step.  The tables hold pointers into the requests' arenas: the requests must outlive
 * the group, and the group's work must be complete before a request is destroyed. */
typedef struct vcl_group vcl_group;   /* opaque */

/* Hash of the kernels, workgroup shapes and argument sizes of f's step (records the step's
 * launches; runs nothing). */
int vcl_fit_step_signature(vcl_fit* f, uint64_t* sig);
/* Device bytes of the work tables of a group of these requests; 0 if they cannot form one. */
size_t vcl_group_table_bytes(vcl_fit* const* fits, int n);
/* Group of n distinct requests with its work tables in caller device memory (uploaded on
 * `stream`; the memory stays the caller's and must not be reused before the group's last step
 * has completed).  VCL_ERR_INVALID if the requests' steps differ in their kernels or shapes. */
int vcl_group_create_in(vcl_fit* const* fits, int n, void* table, size_t table_bytes, vcl_stream_t stream, vcl_group** out);
/* The same with the tables in memory the group allocates (and frees in vcl_group_destroy). */
int vcl_group_create(vcl_fit* const* fits, int n, vcl_stream_t stream, vcl_group** out);
/* One fwd+bwd+Adam step of every request of the group on `stream`. */
int vcl_group_step(vcl_group* g, vcl_stream_t stream);
/* The same step replayed from a HIP graph (captured on the first call, on a stream of the
 * group's own, ordered with `stream` by events). */
int vcl_group_step_graph(vcl_group* g, vcl_stream_t stream);
int vcl_group_size(const vcl_group* g);       /* requests */
int vcl_group_launches(const vcl_group* g);   /* launches per group step */
/* Releases the group (waits for its graph stream; eager steps on a caller stream must have
 * completed). */
int vcl_group_destroy(vcl_group* g);

/* ---------------------------------------------------------------- Serving
 * Continuous batching of fit requests on one GPU (no counterpart in the reference, which runs
 * one preset in a GLUT loop, main.cpp:39-49).  Requests queue with an arrival time on the
 * server clock; each scheduling round retires completed steps, finishes requests that ran all
 * their iterations, admits arrived requests in FIFO order while a slot is free and the
 * request's arena fits the HBM budget, and queues the next steps of every active request on
 * its own stream, so the steps of different requests overlap on the device.  State stays
 * resident in the request's arena until it is released. */
typedef struct vcl_server vcl_server;  /* opaque */

enum {
    VCL_SCHED_FIFO = 0,   /* admit in arrival order; every active request steps */
    VCL_SCHED_SRPT = 1    /* admit the arrived request with the least work first; only the
                             max_stepping active requests with the least remaining work step */
};

typedef struct vcl_server_config {
    size_t hbm_budget;    /* bytes of request arenas the server may hold */
    int max_active;       /* requests admitted (resident) at once */
    int use_graph;        /* 1: replay each request's step from its HIP graph */
    int steps_per_round;  /* steps queued per request per scheduling round (>= 1) */
    int auto_release;     /* 1: free a request's arena as soon as it finishes */
    int policy;           /* VCL_SCHED_FIFO or VCL_SCHED_SRPT */
    int max_stepping;     /* requests stepping at once (0: all active): under FIFO the earliest
                             admitted, under SRPT those with the least remaining work */
    int autotune_steps;   /* > 0: autotune the first request of each size class (pixels over
                             all views in buckets of 4x, triangles in buckets of 8x) with this
                             many timed steps per candidate, and reuse its tuning for the class;
                             0: no tuning */
    int grouped;          /* 1: the requests stepping in a round run in shared launches
                             (vcl_group), one group per step signature and pixel-count bucket
                             (4x), each on its own stream; 0: each request steps on its own
                             stream */
    int group_max_pixels; /* grouped mode: requests with more pixels (over all views) step on
                             their own stream (0: no limit) */
} vcl_server_config;

typedef struct vcl_request_stats {
    uint64_t ticket;
    double arrival_s, admit_s, finish_s;   /* server clock, seconds */
    int iterations;                        /* steps completed */
    float final_loss;
    size_t device_bytes;                   /* arena size */
    int status;                            /* VCL_OK or the error that ended it */
} vcl_request_stats;

int vcl_server_create(const vcl_server_config* c, vcl_server** out);
int vcl_server_destroy(vcl_server* s);
/* Restart the server clock at 0 (arrival times are on this clock). */
int vcl_server_reset_clock(vcl_server* s);
double vcl_server_clock(const vcl_server* s);
/* Queue a request of `iterations` fit steps arriving at `arrival_s`.  The host arrays of `d`
 * must stay valid until the request is admitted.  VCL_ERR_OOM if its arena exceeds the budget. */
int vcl_server_submit(vcl_server* s, const vcl_fit_desc* d, int iterations, double arrival_s, uint64_t* ticket);
/* One non-blocking scheduling round; *pending = requests queued or active. */
int vcl_server_progress(vcl_server* s, int* pending);
/* Scheduling rounds until every submitted request finished. */
int vcl_server_drain(vcl_server* s);
/* Stats of up to `max` finished requests not reported before. */
int vcl_server_finished(vcl_server* s, vcl_request_stats* out, int max, int* n);
/* Copy a named buffer (vcl_fit_buffer names) of a finished, unreleased request to the host. */
int vcl_server_read(vcl_server* s, uint64_t ticket, const char* name, void* host, size_t bytes);
/* Return a finished request's arena to the pool. */
int vcl_server_release(vcl_server* s, uint64_t ticket);
/* Arena bytes held by admitted requests now, at peak, cached for reuse, and the time-average
 * since the clock was reset. */
int vcl_server_memory(vcl_server* s, size_t* in_use, size_t* peak, size_t* cached, double* mean_in_use);
/* The tuning table: size-class keys and their tunings (up to `max`). */
int vcl_server_tunings(vcl_server* s, uint64_t* keys, vcl_tuning* out, int max, int* n);
/* Pre-fill the tuning table (e.g. from a previous server's vcl_server_tunings), so requests of
 * those size classes are not tuned on admission. */
int vcl_server_load_tunings(vcl_server* s, const uint64_t* keys, const vcl_tuning* t, int n);

#ifdef __cplusplus
}
#endif

#endif /* VCL_H_ */
