"""ctypes binding of libvcl_amd.so (the C ABI declared in include/vcl.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvcl_amd.so")

MAXL = 16
vp = C.c_void_p
ci = C.c_int
cf = C.c_float
cd = C.c_double
csz = C.c_size_t

VCL_OK, VCL_ERR_INVALID, VCL_ERR_OOM, VCL_ERR_HIP, VCL_ERR_OVERFLOW = range(5)


class ProjectParams(C.Structure):
    _fields_ = [("vecNum", ci), ("dimention", ci), ("vec", vp), ("mat", vp), ("out", vp), ("views", ci)]


class ProjectGradParams(C.Structure):
    _fields_ = [("vec", vp), ("mat", vp), ("out", vp)]


class RasterizeParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("enableDB", ci), ("xs", cf), ("ys", cf),
                ("proj", vp), ("idx", vp), ("out", vp), ("outDB", vp), ("projNum", ci), ("idxNum", ci),
                ("workspace", vp), ("workspace_size", csz)]


class RasterizeGradParams(C.Structure):
    _fields_ = [("proj", vp), ("out", vp), ("outDB", vp)]


class InterpolateParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("dimention", ci), ("enableDA", ci),
                ("rast", vp), ("rastDB", vp), ("attr", vp), ("idx", vp), ("out", vp), ("outDA", vp), ("attrNum", ci)]


class InterpolateGradParams(C.Structure):
    _fields_ = [("out", vp), ("outDA", vp), ("rast", vp), ("rastDB", vp), ("attr", vp)]


class TexturemapParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("texwidth", ci), ("texheight", ci), ("channel", ci),
                ("miplevel", ci), ("rast", vp), ("uv", vp), ("uvDA", vp), ("texture", vp * MAXL), ("out", vp)]


class TexturemapGradParams(C.Structure):
    _fields_ = [("out", vp), ("uv", vp), ("uvDA", vp), ("texture", vp * MAXL), ("daMode", ci)]


class Texture(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("channel", ci), ("miplevel", ci), ("texture", vp * MAXL)]


class OptimizeParams(C.Structure):
    _fields_ = [("param", vp), ("grad", vp), ("size", ci), ("width", ci), ("height", ci), ("depth", ci)]


class AdamHyperParams(C.Structure):
    _fields_ = [("rhom", cd), ("rhov", cd), ("rhomt", cd), ("rhovt", cd), ("eta", cd), ("eps", cd), ("m", vp), ("v", vp)]


class AntialiasParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("channel", ci), ("s", cf * 2), ("proj", vp), ("idx", vp),
                ("rast", vp), ("in_", vp), ("out", vp), ("projNum", ci)]


class AntialiasGradParams(C.Structure):
    _fields_ = [("proj", vp), ("in_", vp), ("out", vp)]


class NormalAxisParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("rast", vp), ("rot", vp), ("normal", vp), ("normalidx", vp),
                ("pos", vp), ("posidx", vp), ("texel", vp), ("texelidx", vp), ("normalmap", vp), ("out", vp)]


class ViewAxisParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("rast", vp), ("rot", vp), ("view", vp), ("projection", vp),
                ("pvinv", vp), ("out", vp)]


class SGSpecularParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("rast", vp), ("normal", vp), ("view", vp), ("roughness", vp),
                ("ior", cf), ("outAxis", vp), ("outSharpness", vp), ("outAmplitude", vp)]


class SGSpecularGradParams(C.Structure):
    _fields_ = [("normal", vp), ("view", vp), ("roughness", vp), ("ior", cf), ("outAxis", vp), ("outSharpness", vp),
                ("outAmplitude", vp)]


class SphericalGaussianParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("channel", ci), ("rast", vp), ("sgnum", ci), ("axis", vp),
                ("sharpness", vp), ("amplitude", vp), ("normal", vp), ("view", vp), ("diffuse", vp), ("specAxis", vp),
                ("specSharpness", vp), ("specAmplitude", vp), ("out", vp), ("outDiffenv", vp), ("outSpecenv", vp)]


class SphericalGaussianGradParams(C.Structure):
    _fields_ = [("normal", vp), ("diffuse", vp), ("axis", vp), ("sharpness", vp), ("amplitude", vp), ("specAxis", vp),
                ("specSharpness", vp), ("specAmplitude", vp), ("out", vp), ("workspace", vp), ("workspace_size", csz)]


class LossParams(C.Structure):
    _fields_ = [("predict", vp), ("target", vp), ("grad", vp), ("buffer", vp), ("width", ci), ("height", ci), ("depth", ci),
                ("size", ci), ("loss", vp)]


class RotationAdam(C.Structure):
    _fields_ = [("it", ci), ("rhom", cd), ("rhov", cd), ("rhomt", cd), ("rhovt", cd), ("eta", cd), ("eps", cd),
                ("m", C.c_float * 4), ("v", C.c_float * 4)]


class SgdHyperParams(C.Structure):
    _fields_ = [("eta", cd)]


class NadamHyperParams(C.Structure):
    _fields_ = [(n, cd) for n in ("mupow", "mupowt", "mu", "mu0", "mu1", "rho", "alpha", "mut0", "mut1", "rhot", "eps")] + [
        ("m", vp), ("v", vp)]


class NoiseParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("channel", ci), ("intensity", cf), ("in_", vp), ("out", vp)]


class FilterParams(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("depth", ci), ("channel", ci), ("k", ci), ("sigma", vp), ("filter", vp), ("in_", vp),
                ("buf", vp), ("out", vp)]


class FilterGradParams(C.Structure):
    _fields_ = [("in_", vp), ("sigma", vp), ("filter", vp), ("buf", vp), ("bufx", vp), ("bufy", vp), ("out", vp)]


class ComputeNormalParams(C.Structure):
    _fields_ = [("posNum", ci), ("pos", vp), ("vao", vp), ("vaoNum", ci), ("mat", vp), ("out", vp), ("sum", vp), ("angleMode", ci)]


class ComputeNormalGradParams(C.Structure):
    _fields_ = [("pos", vp), ("out", vp)]


class ObjMesh(C.Structure):
    _fields_ = [("posNum", ci), ("texelNum", ci), ("normalNum", ci), ("indexNum", ci), ("pos", C.POINTER(cf)), ("texel", C.POINTER(cf)),
                ("normal", C.POINTER(cf)), ("posIndex", C.POINTER(C.c_uint32)), ("texelIndex", C.POINTER(C.c_uint32)), ("normalIndex", C.POINTER(C.c_uint32))]


class Image(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("channel", ci), ("data", C.POINTER(cf))]


class FitDesc(C.Structure):
    _fields_ = [("width", ci), ("height", ci), ("views", ci), ("V", ci), ("T", ci), ("texwidth", ci), ("texheight", ci),
                ("channel", ci), ("miplevel", ci), ("sgnum", ci),
                ("pos", vp), ("normals", vp), ("uv", vp), ("idx", vp), ("texture", vp), ("axis", vp), ("sharpness", vp),
                ("amplitude", vp), ("mats", vp), ("pvinv", vp), ("rot", vp), ("target", vp),
                ("roughness", cf), ("ior", cf), ("lr", cd), ("beta1", cd), ("beta2", cd), ("eps", cd),
                ("fold_quarter", ci), ("op_by_op", ci)]


class Tuning(C.Structure):
    _fields_ = [("raster_tile", ci), ("bin_chunk", ci), ("splat_rows", ci)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", csz), ("max_active", ci), ("use_graph", ci), ("steps_per_round", ci), ("auto_release", ci),
                ("policy", ci), ("max_stepping", ci), ("autotune_steps", ci), ("grouped", ci),
                ("group_max_pixels", ci)]


class RequestStats(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("arrival_s", cd), ("admit_s", cd), ("finish_s", cd), ("iterations", ci),
                ("final_loss", cf), ("device_bytes", csz), ("status", ci)]


# Every entry point declared in include/vcl.h: name -> (argtypes, restype)
P_ = C.POINTER
SIGNATURES = {
    "vcl_status_string": ([ci], C.c_char_p),
    "vcl_build_info": ([], C.c_char_p),
    "vcl_project_forward": ([P_(ProjectParams), vp], ci),
    "vcl_project_backward": ([P_(ProjectParams), P_(ProjectGradParams), vp], ci),
    "vcl_rasterize_workspace_size": ([ci, ci, ci, ci, ci], csz),
    "vcl_rasterize_forward": ([P_(RasterizeParams), vp], ci),
    "vcl_rasterize_backward": ([P_(RasterizeParams), P_(RasterizeGradParams), vp], ci),
    "vcl_rasterize_status": ([P_(RasterizeParams), vp], ci),
    "vcl_rasterize_set_bin_capacity": ([csz], None),
    "vcl_interpolate_forward": ([P_(InterpolateParams), vp], ci),
    "vcl_interpolate_backward": ([P_(InterpolateParams), P_(InterpolateGradParams), vp], ci),
    "vcl_texturemap_forward": ([P_(TexturemapParams), vp], ci),
    "vcl_texturemap_backward": ([P_(TexturemapParams), P_(TexturemapGradParams), vp], ci),
    "vcl_texture_mip_levels": ([ci, ci, ci], ci),
    "vcl_texture_build_mip": ([P_(Texture), vp], ci),
    "vcl_texture_clamp": ([P_(Texture), cf, cf, vp], ci),
    "vcl_texture_grad_sumup": ([P_(Texture), ci, vp], ci),
    "vcl_texture_grad_clear": ([P_(Texture), vp], ci),
    "vcl_adam_step": ([P_(OptimizeParams), P_(AdamHyperParams), vp], ci),
    "vcl_adam_step_multi": ([ci, P_(OptimizeParams), P_(AdamHyperParams), vp], ci),
    "vcl_texture_adam_update": ([P_(Texture), P_(Texture), P_(AdamHyperParams), cf, cf, ci, vp], ci),
    "vcl_antialias_forward": ([P_(AntialiasParams), vp], ci),
    "vcl_antialias_backward": ([P_(AntialiasParams), P_(AntialiasGradParams), vp], ci),
    "vcl_normal_axis_forward": ([P_(NormalAxisParams), vp], ci),
    "vcl_view_axis_forward": ([P_(ViewAxisParams), vp], ci),
    "vcl_sg_specular_forward": ([P_(SGSpecularParams), vp], ci),
    "vcl_sg_specular_backward": ([P_(SGSpecularParams), P_(SGSpecularGradParams), vp], ci),
    "vcl_spherical_gaussian_workspace_size": ([ci, ci, ci, ci, ci], csz),
    "vcl_spherical_gaussian_forward": ([P_(SphericalGaussianParams), vp], ci),
    "vcl_spherical_gaussian_backward": ([P_(SphericalGaussianParams), P_(SphericalGaussianGradParams), vp], ci),
    "vcl_loss_workspace_size": ([ci], csz),
    "vcl_mse_loss_backward": ([P_(LossParams), vp], ci),
    "vcl_mse_loss_noalloc_max_size": ([], csz),
    "vcl_mse_loss_backward_noalloc": ([P_(LossParams), vp], ci),
    "vcl_ssim_workspace_size": ([ci, ci, ci], csz),
    "vcl_ssim_loss_backward": ([P_(LossParams), ci, vp], ci),
    "vcl_texture_loss_workspace_size": ([ci], csz),
    "vcl_mse_texture_loss": ([P_(LossParams), vp], ci),
    "vcl_sgd_step": ([P_(OptimizeParams), P_(SgdHyperParams), vp], ci),
    "vcl_nadam_step": ([P_(OptimizeParams), P_(NadamHyperParams), vp], ci),
    "vcl_optimize_random_params": ([P_(OptimizeParams), C.c_uint32, cf, cf, vp], ci),
    "vcl_optimize_clamp_params": ([P_(OptimizeParams), cf, cf, vp], ci),
    "vcl_buffer_linear": ([vp, vp, csz, cf, cf, vp], ci),
    "vcl_buffer_add_random": ([vp, csz, cf, cf, C.c_uint32, vp], ci),
    "vcl_buffer_clamp": ([vp, vp, csz, cf, cf, vp], ci),
    "vcl_buffer_step": ([vp, vp, csz, cf, vp], ci),
    "vcl_buffer_normalize": ([vp, vp, ci, ci, vp], ci),
    "vcl_buffer_set_value": ([vp, ci, ci, vp, vp], ci),
    "vcl_buffer_spherical_random": ([vp, ci, C.c_uint32, vp], ci),
    "vcl_attribute_distance_error": ([vp, vp, ci, ci, vp, vp], ci),
    "vcl_texture_bilinear_downsampling": ([P_(Texture), vp], ci),
    "vcl_sg_bake": ([vp, vp, vp, ci, ci, P_(Texture), vp], ci),
    "vcl_sg_normalize": ([vp, vp, vp, ci, ci, vp], ci),
    "vcl_sg_disperse": ([vp, vp, vp, ci, ci, vp, vp], ci),
    "vcl_noise_forward": ([P_(NoiseParams), C.c_uint32, vp], ci),
    "vcl_filter_forward": ([P_(FilterParams), vp], ci),
    "vcl_filter_backward": ([P_(FilterParams), P_(FilterGradParams), vp], ci),
    "vcl_compute_normal_forward": ([P_(ComputeNormalParams), vp], ci),
    "vcl_compute_normal_backward": ([P_(ComputeNormalParams), P_(ComputeNormalGradParams), vp], ci),
    "vcl_rasterize_wireframe": ([ci, ci, ci, vp, vp, vp, vp], ci),
    "vcl_rasterize_idhash": ([ci, ci, ci, vp, vp, vp], ci),
    "vcl_load_obj": ([C.c_char_p, P_(ObjMesh)], ci),
    "vcl_obj_free": ([P_(ObjMesh)], None),
    "vcl_load_bmp": ([C.c_char_p, P_(Image)], ci),
    "vcl_image_free": ([P_(Image)], None),
    "vcl_save_bmp": ([C.c_char_p, vp, ci, ci, ci], ci),
    "vcl_save_pfm": ([C.c_char_p, vp, ci, ci, ci], ci),
    "vcl_load_sg_txt": ([C.c_char_p, ci, vp, vp, vp], ci),
    "vcl_rotation_matrix": ([vp, vp], ci),
    "vcl_view_matrix": ([vp, vp, vp], ci),
    "vcl_look_at_quat": ([vp, vp, vp, vp], ci),
    "vcl_projection_matrix": ([cf, cf, cf, cf, vp], ci),
    "vcl_camera_matrices": ([vp, vp, cf, cf, cf, cf, vp, vp, vp, vp], ci),
    "vcl_camera_backward_host": ([vp, vp, vp, vp, ci, vp, vp, vp], ci),
    "vcl_rotation_backward_host": ([vp, vp, ci, vp], ci),
    "vcl_rotation_adam_step": ([P_(RotationAdam), vp, vp], ci),
    "vcl_fit_create": ([P_(FitDesc), P_(vp)], ci),
    "vcl_fit_arena_bytes": ([P_(FitDesc)], csz),
    "vcl_fit_create_in": ([P_(FitDesc), vp, csz, vp, P_(vp)], ci),
    "vcl_fit_step_signature": ([vp, P_(C.c_uint64)], ci),
    "vcl_group_table_bytes": ([P_(vp), ci], csz),
    "vcl_group_create_in": ([P_(vp), ci, vp, csz, vp, P_(vp)], ci),
    "vcl_group_create": ([P_(vp), ci, vp, P_(vp)], ci),
    "vcl_group_step": ([vp, vp], ci),
    "vcl_group_step_graph": ([vp, vp], ci),
    "vcl_group_size": ([vp], ci),
    "vcl_group_launches": ([vp], ci),
    "vcl_group_destroy": ([vp], ci),
    "vcl_server_create": ([P_(ServerConfig), P_(vp)], ci),
    "vcl_server_destroy": ([vp], ci),
    "vcl_server_reset_clock": ([vp], ci),
    "vcl_server_clock": ([vp], cd),
    "vcl_server_submit": ([vp, P_(FitDesc), ci, cd, P_(C.c_uint64)], ci),
    "vcl_server_progress": ([vp, P_(ci)], ci),
    "vcl_server_drain": ([vp], ci),
    "vcl_server_finished": ([vp, P_(RequestStats), ci, P_(ci)], ci),
    "vcl_server_read": ([vp, C.c_uint64, C.c_char_p, vp, csz], ci),
    "vcl_server_release": ([vp, C.c_uint64], ci),
    "vcl_server_memory": ([vp, P_(csz), P_(csz), P_(csz), P_(cd)], ci),
    "vcl_server_tunings": ([vp, P_(C.c_uint64), P_(Tuning), ci, P_(ci)], ci),
    "vcl_server_load_tunings": ([vp, P_(C.c_uint64), P_(Tuning), ci], ci),
    "vcl_set_tuning": ([P_(Tuning)], ci),
    "vcl_get_tuning": ([P_(Tuning)], ci),
    "vcl_fit_set_tuning": ([vp, P_(Tuning)], ci),
    "vcl_fit_get_tuning": ([vp, P_(Tuning)], ci),
    "vcl_fit_autotune": ([vp, ci, P_(Tuning), P_(cd)], ci),
    "vcl_fit_autotune_scratch_bytes": ([vp], csz),
    "vcl_fit_autotune_in": ([vp, ci, vp, csz, vp, P_(Tuning), P_(cd)], ci),
    "vcl_fit_destroy": ([vp], ci),
    "vcl_fit_set_target_from_render": ([vp, vp], ci),
    "vcl_fit_step": ([vp, vp], ci),
    "vcl_fit_step_graph": ([vp, vp], ci),
    "vcl_fit_buffer": ([vp, C.c_char_p], vp),
    "vcl_fit_device_bytes": ([vp], csz),
    "vcl_fit_read": ([vp, C.c_char_p, vp, csz, vp], ci),
    "vcl_fit_write": ([vp, C.c_char_p, vp, csz, vp], ci),
    "vcl_fit_profile_begin": ([vp, ci, C.c_char_p], ci),
    "vcl_fit_profile_end": ([vp, ci, P_(ci), vp, ci, P_(cd), P_(ci)], ci),
}

_lib = None


def load() -> C.CDLL:
    """Load libvcl_amd.so; raises if it has not been built (python -m vcl_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vcl_amd.build` (no CPU fallback exists)")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


class VclError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status != VCL_OK:
        msg = load().vcl_status_string(status).decode()
        raise VclError(f"{what}: {msg} ({status})")
