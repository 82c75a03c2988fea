"""ctypes binding of the CPU oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this package;
the GPU product path never does.  Arrays are numpy, C-contiguous, float32 or float64 (the
suffix of each entry point picks the type), indices uint32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvcl_oracle.so")


def build(force: bool = False) -> str:
    srcs = [os.path.join(_HERE, f) for f in ("vcl_oracle.cpp", "vcl_oracle_extra.cpp", "oracle_common.hpp")]
    if force or not os.path.exists(LIB_PATH) or any(os.path.getmtime(LIB_PATH) < os.path.getmtime(f) for f in srcs):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


vp, ci, cd, csz = C.c_void_p, C.c_int, C.c_double, C.c_size_t


def _declare(L):
    for s, ct in (("f32", C.c_float), ("f64", C.c_double)):
        sig = {
            "project_fwd": [ci, ci, vp, vp, vp],
            "project_bwd": [ci, ci, vp, vp, vp],
            "raster_fwd": [ci, ci, ci, ci, vp, ci, vp, vp, vp],
            "raster_bwd": [ci, ci, ci, ci, vp, vp, vp, vp, vp],
            "interp_fwd": [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp],
            "interp_bwd": [ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp],
            "tex_fwd": [ci, ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, vp],
            "tex_bwd": [ci, ci, ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, ci],
            "mip_build": [ci, ci, ci, ci, vp],
            "mip_fold": [ci, ci, ci, ci, vp, ci],
            "normal_axis": [ci, ci, ci, vp, vp, vp, vp, vp],
            "view_axis": [ci, ci, ci, vp, vp],
            "sg_spec_fwd": [ci, ci, ci, vp, vp, vp, vp, ct, vp, vp, vp],
            "sg_spec_bwd": [ci, ci, ci, vp, vp, vp, vp, ct, vp, vp, vp],
            "sg_fwd": [ci, ci, ci, ci, ci] + [vp] * 12,
            "sg_bwd": [ci, ci, ci, ci, ci] + [vp] * 18,
            "aa_fwd": [ci, ci, ci, ci, ci, vp, vp, vp, vp, vp],
            "aa_bwd": [ci, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp],
            "mse": [csz, vp, vp, vp],
            "adam": [csz, vp, vp, vp, vp, cd, cd, cd, cd, cd, cd],
            "clamp": [csz, vp, ct, ct],
        }
        for name, args in sig.items():
            f = getattr(L, f"orc_{name}_{s}")
            f.argtypes = args
            f.restype = cd if name == "mse" else None
    for s_, ct in (("f32", C.c_float), ("f64", C.c_double)):
        extra = {
            "ssim": ([ci, ci, ci, vp, vp, vp, vp, ci], cd),
            "filter_fwd": ([ci, ci, ci, ci, ct, vp, vp], None),
            "filter_bwd": ([ci, ci, ci, ci, ct, vp, vp, vp], cd),
            "compute_normal_fwd": ([ci, vp, ci, vp, ci, vp], None),
            "compute_normal_bwd": ([ci, vp, ci, vp, ci, vp, vp], None),
        }
        for name, (args, res) in extra.items():
            f = getattr(L, f"orc_{name}_{s_}")
            f.argtypes = args
            f.restype = res
    u32 = C.c_uint32
    cf = C.c_float
    for name, args, res in (
            ("orc_uniform", [csz, vp, u32, u32, vp], None),
            ("orc_noise", [ci, ci, ci, ci, cf, vp, vp, u32], None),
            ("orc_sgd", [csz, vp, vp, cd], None),
            ("orc_nadam", [csz, vp, vp, vp, vp] + [cd] * 9, None),
            ("orc_random_params", [csz, vp, u32, cf, cf], None),
            ("orc_linear", [csz, vp, vp, cf, cf], None),
            ("orc_add_random", [csz, vp, cf, cf, u32], None),
            ("orc_clamp_buf", [csz, vp, vp, cf, cf], None),
            ("orc_step_buf", [csz, vp, vp, cf], None),
            ("orc_normalize_rows", [ci, ci, vp, vp], None),
            ("orc_spherical_random", [ci, vp, u32], None),
            ("orc_distance_error", [ci, ci, vp, vp], cd),
            ("orc_bilinear_downsampling", [ci, ci, ci, ci, vp], None),
            ("orc_sg_bake", [ci, ci, vp, vp, vp, ci, ci, vp], None),
            ("orc_sg_normalize", [ci, ci, vp, vp, vp], None),
            ("orc_sg_disperse", [ci, ci, vp, vp, vp, vp], None),
            ("orc_texture_loss", [ci, ci, ci, vp, vp, vp], cd),
            ("orc_raster_idhash", [csz, vp, vp], None),
            ("orc_raster_wireframe", [csz, vp, vp, vp], None)):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    L.orc_raster_brute.argtypes = [ci, ci, vp, ci, vp, vp]
    L.orc_raster_coverage_count.argtypes = [ci, ci, vp, ci, vp, vp]
    L.orc_mip_levels.argtypes = [ci, ci, ci]
    L.orc_mip_levels.restype = ci
    L.orc_set_threads.argtypes = [ci]
    L.orc_work_create.restype = vp
    L.orc_work_destroy.argtypes = [vp]
    L.orc_fit_step_f32.argtypes = [vp, vp]
    L.orc_work_buffer.argtypes = [vp, C.c_char_p]
    L.orc_work_buffer.restype = C.POINTER(C.c_float)
    L.orc_work64_create.restype = vp
    L.orc_work64_destroy.argtypes = [vp]
    L.orc_fit_grads_f64.argtypes = [vp, vp]
    L.orc_fit_grads_f64.restype = cd
    L.orc_work64_buffer.argtypes = [vp, C.c_char_p]
    L.orc_work64_buffer.restype = C.POINTER(C.c_double)


def P(a):
    """Pointer of a C-contiguous numpy array (or None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def sfx(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def set_threads(n: int):
    lib().orc_set_threads(int(n))


class OrcScene(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("W", "H", "V", "T", "C", "K", "tw", "th", "L")] + [
        ("pos", vp), ("normals", vp), ("uv", vp), ("idx", vp), ("mat", vp), ("rot", vp), ("pvinv", vp),
        ("tex", vp), ("axis", vp), ("sharp", vp), ("amp", vp), ("rough", vp), ("ior", C.c_float),
        ("target", vp), ("fold_quarter", C.c_int),
        ("m_pos", vp), ("v_pos", vp), ("m_tex", vp), ("v_tex", vp), ("m_axis", vp), ("v_axis", vp),
        ("m_sharp", vp), ("v_sharp", vp), ("m_amp", vp), ("v_amp", vp),
        ("lr", cd), ("b1", cd), ("b2", cd), ("eps", cd), ("b1t", cd), ("b2t", cd),
        ("out_image", vp), ("loss", cd), ("views", C.c_int)]
