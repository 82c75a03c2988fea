// internal.hpp -- entry points shared between the library's translation units (not part of
// the public C ABI in include/vcl.h).
#pragma once
#include "common.hpp"
#include "group.hpp"

// Process-wide kernel tuning (vcl_set_tuning); read by the launch code of the raster and
// texture-backward entry points.
extern vcl_tuning g_vcl_tuning;

// The last pass of a C == 4 mip rebuild (levels l .. l + nl - 1 from the level above, of
// sw x sh texels), when it is a single workgroup's.
struct VclMipTail {
    int valid;
    const float* src;
    int sw, sh, nl;
    float* dst[6];
    float* zg[6];
};

extern "C" {
// ViewAxis forward with the per-view inverse matrices already on the device (depth x 9 floats).
int vcl_view_axis_forward_dev(int W, int H, int depth, const float* pvinv_dev, float* out, vcl_stream_t stream);
// Interpolate (uv: 2 attributes with derivatives) and the C == 4 texture forward in one launch;
// tp->uv and tp->uvDA are ip's outputs.
// zero (may be null): a 4-float-per-pixel buffer cleared in the same pass.
int vcl_interp_tex_dev(const vcl_interpolate_params* ip, const vcl_texturemap_params* tp, float* zero, vcl_stream_t stream);
// vcl_project_forward that also clears zero (shaped like p->out, may be null), when bt is not
// null applies the step-start bump bt[0] *= rhom, bt[1] *= rhov once, and, when rc is not null,
// does the rasterizer's workspace reset that vcl_rasterize_clear_spec describes.
int vcl_project_forward_dev(const vcl_project_params* p, float* zero, double* bt, double rhom, double rhov, const WsClear* rc,
                            vcl_stream_t stream);
// The workspace reset of the next vcl_rasterize_forward(_dev) with these params (same tuning).
int vcl_rasterize_clear_spec(const vcl_rasterize_params* p, WsClear* c);
// vcl_rasterize_forward; cleared: the workspace reset was already done (vcl_rasterize_clear_spec).
// keys (may be null): the per-pixel resolve is left to the caller, which gets the winning-key
// buffer (one 64-bit key per pixel, all-ones where no triangle won; raster_resolve.hpp) in *keys
// and must write p->out / p->outDB from it (vcl_pixel_prep_dev does).
int vcl_rasterize_forward_dev(const vcl_rasterize_params* p, int cleared, vcl_stream_t stream, const unsigned long long** keys = nullptr);
// NormalAxis, ViewAxis (inverse projection-view blocks on the device) and SGSpecular forward in one
// launch; sp->normal and sp->view are the outputs of the first two.
int vcl_shade_prep_dev(const vcl_normal_axis_params* np, const float* pvinv_dev, const vcl_sg_specular_params* sp, vcl_stream_t stream);
// The rasterizer's resolve (keys: the winning-key buffer of vcl_rasterize_forward_dev with
// resolve deferred; rp->out and rp->outDB written), uv interpolation + texture sample (C == 4)
// and the shading inputs in one pass: the outputs of k_resolve, vcl_interp_tex_dev and
// vcl_shade_prep_dev (normals indexed like the uv attributes).
int vcl_pixel_prep_dev(const vcl_rasterize_params* rp, const unsigned long long* keys, const vcl_interpolate_params* ip,
                       const vcl_texturemap_params* tp, const vcl_normal_axis_params* np, const float* pvinv_dev, const vcl_sg_specular_params* sp,
                       vcl_stream_t stream);
// Texture backward then interpolate backward (the latter's output grad being the former's uv
// grad): fused into one kernel when the tiled C == 4 texture backward applies, a 2-channel
// attribute feeds the uv and only the rast grad is wanted; otherwise the two ops in turn.
struct vcl_sg_lobe_sums;
// stage (may be null or not valid): an SG segment stage (vcl_sg_lobe_sums) the texture backward's
// launch also runs, as workgroups after its own.
int vcl_texture_interp_backward_dev(const vcl_texturemap_params* tp, const vcl_texturemap_grad_params* tg, const vcl_interpolate_params* ip,
                                    const vcl_interpolate_grad_params* ig, vcl_stream_t stream, const vcl_sg_lobe_sums* stage = nullptr);
// Fused texture update.  part bit 1: fold + Adam + clamp on level 0; bit 2: rebuild the mips and
// clear the coarse grads; bit 4 (C == 4): level 1 and its grad clear move from the mip rebuild
// into the level-0 update (one 2 x 2 quad per thread), so the two calls of a split update must
// both pass it.  bt_dev: device {rhomt, rhovt} used instead of h->rhomt/rhovt (null: the host
// values).
// tail (may be null): a last mip pass of a single workgroup is not launched but described here
// (valid = 1) for the caller's Adam launch to run (vcl_adam_step_multi_proj_dev).
int vcl_texture_adam_update_parts(const vcl_texture* tex, const vcl_texture* grad, const vcl_adam_hyper_params* h, float lo, float hi,
                                  int quarter, int part, const double* bt_dev, vcl_stream_t stream, VclMipTail* tail = nullptr);
// vcl_adam_step_multi with the bias-correction powers read from the device (bt_dev, may be null).
// zero_grad: clear each grad after it is used.
int vcl_adam_step_multi_dev(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, const double* bt_dev, int zero_grad,
                            vcl_stream_t stream);
// The same with the project backward folded in: tensor 0 is proj's positions (3 floats per
// vertex), and its grad, before the update, also takes the vec grad of vcl_project_backward
// with proj->out as the output grad (clip-space, 4 floats per vertex and view).
// The SG backward's per-lobe sums left for the caller: seg holds nseg rows of K x 8 columns
// (axis 3, sharpness 1, amplitude up to 4), summed over the rows in order into the lobe grads.
// Segment stage of those sums (seg = column sums over consecutive groups of seg_rows rows of the
// rows x cols partials in part), when the caller runs it (defer_stage).
struct VclColsumStage {
    const float* part;
    int rows, cols, seg_rows;
    float* seg;
};
struct vcl_sg_lobe_sums {
    const float* seg;
    int nseg, K, C;
    int defer_stage;       // in: leave the segment stage to the caller (stage set, valid = 1)
    int stage_valid;
    VclColsumStage stage;
};
// vcl_spherical_gaussian_backward that, when lobes is not null and all three lobe grads are
// requested, leaves the lobe sums to the caller (lobes->seg set) instead of adding them, and with
// lobes->defer_stage also the segment stage that fills seg (lobes->stage).
// gd_fresh: g->diffuse holds nothing yet (the pass writes it as if it had been cleared, so the
// caller needs no clear of its own).
int vcl_spherical_gaussian_backward_dev(const vcl_spherical_gaussian_params* p, const vcl_spherical_gaussian_grad_params* g,
                                        vcl_sg_lobe_sums* lobes, int gd_fresh, vcl_stream_t stream);
// loss (may be null): the launch also sums the loss partials that vcl_mse_loss_grad_dev left in
// loss->buffer into *loss->loss (one more workgroup; n <= 8 tensors).  lobes (may be null, or
// with seg null): tensors 1..3 are the lobe axis, sharpness and amplitude, whose grads first
// take those sums.
// loss_parts >= 0: the number of partials in loss->buffer (else one per 1,024 elements, k_mse's).
// mip_tail (may be null or not valid): a last mip pass the launch also runs (one workgroup more).
int vcl_adam_step_multi_proj_dev(int n, const vcl_optimize_params* p, const vcl_adam_hyper_params* h, const double* bt_dev, int zero_grad,
                                 const vcl_project_params* proj, const vcl_loss_params* loss, const vcl_sg_lobe_sums* lobes,
                                 vcl_stream_t stream, long long loss_parts = -1, const VclMipTail* mip_tail = nullptr);
// vcl_antialias_forward that, given target, gout and part, also writes the L2 loss's grad
// out - target and one sum of squared grads per 16 x 16 tile (vcl_antialias_loss_parts of them).
int vcl_antialias_forward_dev(const vcl_antialias_params* p, const float* target, float* gout, float* part, vcl_stream_t stream);
size_t vcl_antialias_loss_parts(int width, int height, int depth);
// vcl_mse_loss_backward without the final sum: the grad and the per-block partials only.
int vcl_mse_loss_grad_dev(const vcl_loss_params* p, vcl_stream_t stream);
// bt_dev[0] *= rhom; bt_dev[1] *= rhov (one thread; the step-start bump of Adam::step).
}

// Shared launches over several requests (group.hip): the launches of one step of F, recorded
// (nothing runs); VCL_ERR_INVALID if the step holds device work that cannot be grouped.
int vcl_fit_record_step(vcl_fit* F, VRecorder* rec);
// Adds k to F's step count (steps taken through a group).
void vcl_fit_count_steps(vcl_fit* F, int k);
