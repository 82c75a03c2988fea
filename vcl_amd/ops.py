"""Python mirror of the reference op surface over the C ABI (torch tensors as device memory).

Each function launches the HIP kernels of libvcl_amd.so on torch's current stream; torch is
only the allocator and stream provider.  Names follow the reference layers
(Project/Rasterize/Interpolate/Texturemap/Antialias/NormalAxis/ViewAxis/SGSpecular/
SphericalGaussian/MSELoss/Adam, nvdiffrast/nvdiffrast/*.h) and keep their argument meaning;
errors raise VclError (the reference prints and exits).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

F32 = torch.float32


def _p(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _lib():
    return L.load()


def empty(*shape, dtype=F32):
    return torch.empty(*shape, dtype=dtype, device="cuda")


def zeros(*shape, dtype=F32):
    return torch.zeros(*shape, dtype=dtype, device="cuda")


# ---------------------------------------------------------------- Project (project.h)
def project(pos, mats, views=1):
    V = pos.shape[0]
    out = empty(views, V, 4)
    p = L.ProjectParams(V, 4, _p(pos).value, _p(mats).value, _p(out).value, views)
    L.check(_lib().vcl_project_forward(C.byref(p), _s()), "project")
    return out


def project_backward(pos, mats, gclip, gpos=None, gmat=None, views=1):
    p = L.ProjectParams(pos.shape[0], 4, _p(pos).value, _p(mats).value, None, views)
    g = L.ProjectGradParams(_p(gpos), _p(gmat), _p(gclip))
    L.check(_lib().vcl_project_backward(C.byref(p), C.byref(g), _s()), "project_backward")


# ---------------------------------------------------------------- Rasterize (rasterize.h)
class Rasterizer:
    """Holds the rasterizer workspace for one (W, H, depth, V, T) shape."""

    def __init__(self, W, H, depth, V, T):
        self.W, self.H, self.depth, self.V, self.T = W, H, depth, V, T
        n = _lib().vcl_rasterize_workspace_size(W, H, depth, V, T)
        self.ws = torch.empty(max(int(n), 4), dtype=torch.uint8, device="cuda")

    def params(self, clip, idx, out, db):
        return L.RasterizeParams(self.W, self.H, self.depth, 1 if db is not None else 0, 2.0 / self.W, 2.0 / self.H,
                                 _p(clip).value, _p(idx).value, _p(out).value, _p(db).value if db is not None else None,
                                 self.V, self.T, _p(self.ws).value, self.ws.numel())

    def forward(self, clip, idx, enable_db=True):
        out = empty(self.depth, self.H, self.W, 4)
        db = empty(self.depth, self.H, self.W, 4) if enable_db else None
        p = self.params(clip, idx, out, db)
        L.check(_lib().vcl_rasterize_forward(C.byref(p), _s()), "rasterize")
        return out, db

    def status(self, clip, idx, out, db):
        p = self.params(clip, idx, out, db)
        return _lib().vcl_rasterize_status(C.byref(p), _s())

    def backward(self, clip, idx, rast, grast, gclip):
        p = self.params(clip, idx, rast, None)
        g = L.RasterizeGradParams(_p(gclip), _p(grast), None)
        L.check(_lib().vcl_rasterize_backward(C.byref(p), C.byref(g), _s()), "rasterize_backward")


def rasterize(clip, idx, W, H, depth=1, enable_db=True):
    r = Rasterizer(W, H, depth, clip.shape[-2], idx.shape[0])
    return r.forward(clip, idx, enable_db)


# ---------------------------------------------------------------- Interpolate (interpolate.h)
def _interp_params(rast, db, attr, idx, out, outDA):
    depth, H, W = rast.shape[0], rast.shape[1], rast.shape[2]
    D = attr.shape[1]
    return L.InterpolateParams(W, H, depth, D, 1 if outDA is not None else 0, _p(rast).value, _p(db).value if db is not None else None,
                               _p(attr).value, _p(idx).value, _p(out).value if out is not None else None,
                               _p(outDA).value if outDA is not None else None, attr.shape[0])


def interpolate(rast, db, attr, idx, enable_da=True):
    depth, H, W = rast.shape[:3]
    D = attr.shape[1]
    out = empty(depth, H, W, D)
    outDA = empty(depth, H, W, D, 2) if enable_da else None
    p = _interp_params(rast, db if enable_da else None, attr, idx, out, outDA)
    L.check(_lib().vcl_interpolate_forward(C.byref(p), _s()), "interpolate")
    return out, outDA


def interpolate_backward(rast, attr, idx, gout, gattr=None, grast=None):
    p = _interp_params(rast, None, attr, idx, None, None)
    g = L.InterpolateGradParams(_p(gout), None, _p(grast), None, _p(gattr))
    L.check(_lib().vcl_interpolate_backward(C.byref(p), C.byref(g), _s()), "interpolate_backward")


# ---------------------------------------------------------------- Texturemap (texturemap.h)
def level_views(flat, tw, th, C_, L_):
    """Split a flat all-levels tensor into per-level (h, w, C) views."""
    out, off = [], 0
    for l in range(L_):
        n = (tw >> l) * (th >> l) * C_
        out.append(flat[off:off + n].view(th >> l, tw >> l, C_))
        off += n
    return out


def _tex_params(rast, uv, uvDA, levels, out):
    depth, H, W = rast.shape[:3]
    th, tw, C_ = levels[0].shape
    tex = (C.c_void_p * L.MAXL)(*([_p(x).value for x in levels] + [None] * (L.MAXL - len(levels))))
    return L.TexturemapParams(W, H, depth, tw, th, C_, len(levels), _p(rast).value, _p(uv).value,
                              _p(uvDA).value if uvDA is not None else None, tex, _p(out).value if out is not None else None)


def texture(rast, uv, uvDA, levels):
    depth, H, W = rast.shape[:3]
    out = empty(depth, H, W, levels[0].shape[2])
    p = _tex_params(rast, uv, uvDA, levels, out)
    L.check(_lib().vcl_texturemap_forward(C.byref(p), _s()), "texture")
    return out


def texture_backward(rast, uv, uvDA, levels, gout, glevels=None, guv=None, guvDA=None, da_mode=0):
    p = _tex_params(rast, uv, uvDA, levels, None)
    gt = (C.c_void_p * L.MAXL)(*([(_p(x).value if x is not None else None) for x in (glevels or [])] + [None] * (L.MAXL - len(glevels or []))))
    g = L.TexturemapGradParams(_p(gout), _p(guv), _p(guvDA), gt, da_mode)
    L.check(_lib().vcl_texturemap_backward(C.byref(p), C.byref(g), _s()), "texture_backward")


def _texture_struct(levels):
    th, tw, C_ = levels[0].shape
    arr = (C.c_void_p * L.MAXL)(*([_p(x).value for x in levels] + [None] * (L.MAXL - len(levels))))
    return L.Texture(tw, th, C_, len(levels), arr)


def build_mip(levels):
    t = _texture_struct(levels)
    L.check(_lib().vcl_texture_build_mip(C.byref(t), _s()), "build_mip")


def grad_sumup(glevels, quarter=True):
    t = _texture_struct(glevels)
    L.check(_lib().vcl_texture_grad_sumup(C.byref(t), 1 if quarter else 0, _s()), "grad_sumup")


def texture_clamp(levels, lo=0.0, hi=1.0):
    t = _texture_struct(levels)
    L.check(_lib().vcl_texture_clamp(C.byref(t), lo, hi, _s()), "texture_clamp")


def texture_adam_update(levels, glevels, m, v, hyper, lo=0.0, hi=1.0, quarter=True):
    t, g = _texture_struct(levels), _texture_struct(glevels)
    h = L.AdamHyperParams(*hyper, _p(m).value, _p(v).value)
    L.check(_lib().vcl_texture_adam_update(C.byref(t), C.byref(g), C.byref(h), lo, hi, 1 if quarter else 0, _s()), "texture_adam_update")


# ---------------------------------------------------------------- Adam (optimize.h)
def adam_step(params, grads, ms, vs, hyper):
    """hyper = (rhom, rhov, rhomt, rhovt, eta, eps): one multi-tensor launch."""
    n = len(params)
    op = (L.OptimizeParams * n)(*[L.OptimizeParams(_p(p).value, _p(g).value, p.numel(), p.numel(), 1, 1) for p, g in zip(params, grads)])
    hp = (L.AdamHyperParams * n)(*[L.AdamHyperParams(*hyper, _p(m).value, _p(v).value) for m, v in zip(ms, vs)])
    L.check(_lib().vcl_adam_step_multi(n, op, hp, _s()), "adam")


# ---------------------------------------------------------------- Material (material.h)
def normal_axis(rast, normals, nidx, rot):
    depth, H, W = rast.shape[:3]
    out = empty(depth, H, W, 3)
    p = L.NormalAxisParams(W, H, depth, _p(rast).value, _p(rot).value, _p(normals).value, _p(nidx).value, None, None, None, None, None,
                           _p(out).value)
    L.check(_lib().vcl_normal_axis_forward(C.byref(p), _s()), "normal_axis")
    return out


def view_axis(W, H, views_np, projs_np):
    """views_np / projs_np: host float32 arrays of depth x 16 (column-major)."""
    depth = views_np.reshape(-1, 16).shape[0]
    out = empty(depth, H, W, 3)
    vh = np.ascontiguousarray(views_np, np.float32).reshape(-1)
    ph = np.ascontiguousarray(projs_np, np.float32).reshape(-1)
    p = L.ViewAxisParams(W, H, depth, None, None, vh.ctypes.data, ph.ctypes.data, None, _p(out).value)
    L.check(_lib().vcl_view_axis_forward(C.byref(p), _s()), "view_axis")
    return out


def sg_specular(rast, normal, view, rough, ior):
    depth, H, W = rast.shape[:3]
    ax, sh, am = empty(depth, H, W, 3), empty(depth, H, W), empty(depth, H, W)
    p = L.SGSpecularParams(W, H, depth, _p(rast).value, _p(normal).value, _p(view).value, _p(rough).value, ior, _p(ax).value,
                           _p(sh).value, _p(am).value)
    L.check(_lib().vcl_sg_specular_forward(C.byref(p), _s()), "sg_specular")
    return ax, sh, am


def sg_specular_backward(rast, normal, view, rough, ior, gAmp, gSharp, grough):
    depth, H, W = rast.shape[:3]
    p = L.SGSpecularParams(W, H, depth, _p(rast).value, _p(normal).value, _p(view).value, _p(rough).value, ior, None, None, None)
    g = L.SGSpecularGradParams(None, None, _p(grough).value, ior, None, _p(gSharp).value, _p(gAmp).value)
    L.check(_lib().vcl_sg_specular_backward(C.byref(p), C.byref(g), _s()), "sg_specular_backward")


def _sg_params(rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out=None, denv=None, senv=None):
    """SphericalGaussianKernelParams; denv (the forward's outDiffenv) enables the lobe-parallel backward."""
    depth, H, W = rast.shape[:3]
    C_ = diffuse.shape[-1]
    return L.SphericalGaussianParams(W, H, depth, C_, _p(rast).value, sharp.numel(), _p(axis).value, _p(sharp).value, _p(amp).value,
                                     _p(normal).value, None, _p(diffuse).value, _p(sAxis).value, _p(sSharp).value, _p(sAmp).value,
                                     _p(out), _p(denv), _p(senv))


def spherical_gaussian(rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp):
    out, denv, senv = torch.empty_like(diffuse), torch.empty_like(diffuse), torch.empty_like(diffuse)
    p = _sg_params(rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, out, denv, senv)
    L.check(_lib().vcl_spherical_gaussian_forward(C.byref(p), _s()), "spherical_gaussian")
    return out, denv, senv


def spherical_gaussian_backward(rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, gout, gdiffuse=None, gaxis=None,
                                gsharp=None, gamp=None, gnormal=None, gsAxis=None, gsSharp=None, gsAmp=None, denv=None):
    depth, H, W = rast.shape[:3]
    p = _sg_params(rast, axis, sharp, amp, normal, diffuse, sAxis, sSharp, sAmp, denv=denv)
    n = _lib().vcl_spherical_gaussian_workspace_size(W, H, depth, diffuse.shape[-1], sharp.numel())
    ws = torch.empty(int(n), dtype=torch.uint8, device="cuda")
    g = L.SphericalGaussianGradParams(_p(gnormal), _p(gdiffuse), _p(gaxis), _p(gsharp), _p(gamp), _p(gsAxis), _p(gsSharp), _p(gsAmp),
                                      _p(gout), _p(ws), ws.numel())
    L.check(_lib().vcl_spherical_gaussian_backward(C.byref(p), C.byref(g), _s()), "spherical_gaussian_backward")


# ---------------------------------------------------------------- Antialias (antialias.h)
def _aa_params(clip, idx, rast, inp, out):
    depth, H, W = rast.shape[:3]
    s = (C.c_float * 2)(W / 2.0, H / 2.0)
    return L.AntialiasParams(W, H, depth, inp.shape[-1], s, _p(clip).value, _p(idx).value, _p(rast).value, _p(inp).value, _p(out),
                             clip.shape[-2])


def antialias(clip, idx, rast, inp):
    out = torch.empty_like(inp)
    p = _aa_params(clip, idx, rast, inp, out)
    L.check(_lib().vcl_antialias_forward(C.byref(p), _s()), "antialias")
    return out


def antialias_backward(clip, idx, rast, inp, gout, gclip=None):
    gin = torch.empty_like(inp)
    p = _aa_params(clip, idx, rast, inp, None)
    g = L.AntialiasGradParams(_p(gclip), _p(gin).value, _p(gout).value)
    L.check(_lib().vcl_antialias_backward(C.byref(p), C.byref(g), _s()), "antialias_backward")
    return gin


# ---------------------------------------------------------------- MSELoss (loss.h)
def mse_loss(pred, target):
    n = pred.numel()
    grad = torch.empty_like(pred)
    ws = torch.empty(int(_lib().vcl_loss_workspace_size(n)), dtype=torch.uint8, device="cuda")
    loss = empty(1)
    p = L.LossParams(_p(pred).value, _p(target).value, _p(grad).value, _p(ws).value, n, 1, 1, n, _p(loss).value)
    L.check(_lib().vcl_mse_loss_backward(C.byref(p), _s()), "mse")
    return grad, loss


# ---------------------------------------------------------------- further losses (loss.h)
def ssim_loss(pred, target, grad_mode=0):
    """SSIMLoss::backward over HWC tensors (H, W, D): returns (grad, loss, 1 - SSIM map).
    grad_mode 0: the reference's gradient formula; 1: the exact gradient of sum(1 - SSIM)."""
    H, W, D = pred.shape
    n = pred.numel()
    grad = torch.empty_like(pred)
    ws = torch.empty(int(_lib().vcl_ssim_workspace_size(W, H, D)), dtype=torch.uint8, device="cuda")
    loss = empty(1)
    p = L.LossParams(_p(pred).value, _p(target).value, _p(grad).value, _p(ws).value, W, H, D, n, _p(loss).value)
    L.check(_lib().vcl_ssim_loss_backward(C.byref(p), int(grad_mode), _s()), "ssim")
    return grad, loss, ws[: 4 * n].view(torch.float32).view(H, W, D)


def texture_loss(pred, target):
    """MSELoss::textureloss over (H, W, D): returns (loss, per-element map)."""
    H, W, D = pred.shape
    n = pred.numel()
    ws = torch.empty(int(_lib().vcl_texture_loss_workspace_size(n)), dtype=torch.uint8, device="cuda")
    loss = empty(1)
    p = L.LossParams(_p(pred).value, _p(target).value, None, _p(ws).value, W, H, D, n, _p(loss).value)
    L.check(_lib().vcl_mse_texture_loss(C.byref(p), _s()), "texture_loss")
    return loss, ws[: 4 * n].view(torch.float32).view(H, W, D)


# ---------------------------------------------------------------- further optimizers (optimize.h)
def _opt(param, grad=None):
    return L.OptimizeParams(_p(param).value, _p(grad).value if grad is not None else None, param.numel(), param.numel(), 1, 1)


def sgd_step(param, grad, eta):
    p = _opt(param, grad)
    L.check(_lib().vcl_sgd_step(C.byref(p), C.byref(L.SgdHyperParams(eta)), _s()), "sgd")


def nadam_step(param, grad, m, v, mu, mu0, mu1, rho, alpha, mut0, mut1, rhot, eps):
    p = _opt(param, grad)
    h = L.NadamHyperParams(0.0, 0.0, mu, mu0, mu1, rho, alpha, mut0, mut1, rhot, eps, _p(m).value, _p(v).value)
    L.check(_lib().vcl_nadam_step(C.byref(p), C.byref(h), _s()), "nadam")


def random_params(param, seed, lo, hi):
    p = _opt(param)
    L.check(_lib().vcl_optimize_random_params(C.byref(p), int(seed), lo, hi, _s()), "random_params")


def clamp_params(param, lo, hi):
    p = _opt(param)
    L.check(_lib().vcl_optimize_clamp_params(C.byref(p), lo, hi, _s()), "clamp_params")


# ---------------------------------------------------------------- buffer utilities (buffer.h)
def buffer_linear(dst, src, w, b):
    L.check(_lib().vcl_buffer_linear(_p(dst), _p(src), dst.numel(), w, b, _s()), "linear")


def buffer_add_random(dst, lo, hi, seed):
    L.check(_lib().vcl_buffer_add_random(_p(dst), dst.numel(), lo, hi, int(seed), _s()), "add_random")


def buffer_clamp(dst, src, lo, hi):
    L.check(_lib().vcl_buffer_clamp(_p(dst), _p(src), dst.numel(), lo, hi, _s()), "clamp")


def buffer_step(dst, src, thr):
    L.check(_lib().vcl_buffer_step(_p(dst), _p(src), dst.numel(), thr, _s()), "step")


def buffer_normalize(dst, src, dim):
    L.check(_lib().vcl_buffer_normalize(_p(dst), _p(src), dst.numel() // dim, dim, _s()), "normalize")


def buffer_set_value(dst, value):
    v = np.ascontiguousarray(value, np.float32)
    L.check(_lib().vcl_buffer_set_value(_p(dst), dst.numel() // v.size, v.size, v.ctypes.data, _s()), "set_value")


def buffer_spherical_random(dst, seed):
    L.check(_lib().vcl_buffer_spherical_random(_p(dst), dst.numel() // 3, int(seed), _s()), "spherical_random")


def distance_error(pred, target):
    out = empty(1)
    L.check(_lib().vcl_attribute_distance_error(_p(pred), _p(target), pred.shape[0], pred.shape[1], _p(out), _s()), "distance_error")
    return out


def bilinear_downsampling(levels):
    t = _texture_struct(levels)
    L.check(_lib().vcl_texture_bilinear_downsampling(C.byref(t), _s()), "bilinear_downsampling")


def sg_bake(axis, sharp, amp, levels):
    t = _texture_struct(levels)
    L.check(_lib().vcl_sg_bake(_p(axis), _p(sharp), _p(amp), sharp.numel(), amp.shape[1], C.byref(t), _s()), "sg_bake")


def sg_normalize(axis, sharp, amp):
    L.check(_lib().vcl_sg_normalize(_p(axis), _p(sharp), _p(amp), sharp.numel(), amp.shape[1], _s()), "sg_normalize")


def sg_disperse(axis, sharp, amp):
    buf = empty(axis.shape)
    L.check(_lib().vcl_sg_disperse(_p(axis), _p(sharp), _p(amp), sharp.numel(), amp.shape[1], _p(buf), _s()), "sg_disperse")
    return buf


# ---------------------------------------------------------------- Noise (extention.h), Filter (filter.h)
def noise(inp, intensity, seed):
    depth, H, W, C_ = inp.shape
    out = torch.empty_like(inp)
    p = L.NoiseParams(W, H, depth, C_, intensity, _p(inp).value, _p(out).value)
    L.check(_lib().vcl_noise_forward(C.byref(p), int(seed), _s()), "noise")
    return out


def _filter_params(inp, sigma_dev, out):
    depth, H, W, C_ = inp.shape
    filt = empty(256)
    buf = torch.empty_like(inp)
    return L.FilterParams(W, H, depth, C_, 0, _p(sigma_dev).value, _p(filt).value, _p(inp).value, _p(buf).value, _p(out).value), (filt, buf)


def filter_forward(inp, sigma_dev):
    out = torch.empty_like(inp)
    p, keep = _filter_params(inp, sigma_dev, out)
    L.check(_lib().vcl_filter_forward(C.byref(p), _s()), "filter")
    return out


def filter_backward(inp, sigma_dev, gout):
    """Returns (grad of the input, d loss / d sigma as a 1-element tensor)."""
    p, keep = _filter_params(inp, sigma_dev, torch.empty_like(inp))
    gin = torch.empty_like(inp)
    gsig = empty(1)
    gf, b0, b1, b2 = empty(256), torch.empty_like(inp), torch.empty_like(inp), torch.empty_like(inp)
    g = L.FilterGradParams(_p(gin).value, _p(gsig).value, _p(gf).value, _p(b0).value, _p(b1).value, _p(b2).value, _p(gout).value)
    L.check(_lib().vcl_filter_backward(C.byref(p), C.byref(g), _s()), "filter_backward")
    torch.cuda.current_stream().synchronize()
    return gin, gsig


# ---------------------------------------------------------------- ComputeNormal (computenormal.h)
def compute_normal(pos, idx, mode=0, gout=None):
    """Vertex normals; with gout also returns the position grad."""
    V = pos.shape[0]
    out = empty(V, 3)
    ssum = empty(V, 3)
    p = L.ComputeNormalParams(V, _p(pos).value, _p(idx).value, idx.shape[0], None, _p(out).value, _p(ssum).value, int(mode))
    L.check(_lib().vcl_compute_normal_forward(C.byref(p), _s()), "compute_normal")
    if gout is None:
        return out
    gpos = zeros(V, 3)
    g = L.ComputeNormalGradParams(_p(gpos).value, _p(gout).value)
    L.check(_lib().vcl_compute_normal_backward(C.byref(p), C.byref(g), _s()), "compute_normal_backward")
    return out, gpos


# ---------------------------------------------------------------- debug raster shadings (rasterize.h)
def rasterize_wireframe(rast, db):
    depth, H, W = rast.shape[:3]
    out = empty(depth, H, W, 4)
    L.check(_lib().vcl_rasterize_wireframe(W, H, depth, _p(rast), _p(db), _p(out), _s()), "wireframe")
    return out


def rasterize_idhash(rast):
    depth, H, W = rast.shape[:3]
    out = empty(depth, H, W, 4)
    L.check(_lib().vcl_rasterize_idhash(W, H, depth, _p(rast), _p(out), _s()), "idhash")
    return out


# ---------------------------------------------------------------- Fit engine
def fit_desc(scene, target=None, lr=1e-2, betas=(0.9, 0.999), eps=1e-8, fold_quarter=True, op_by_op=False, views=None):
    """vcl_fit_desc of a Scene; returns (desc, keep) where `keep` holds the host arrays the
    descriptor points into (they must outlive the upload)."""
    cams = scene.cams if views is None else views
    keep = []

    def h(a, dt=np.float32):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    mats = np.stack([c.mat for c in cams]).astype(np.float32)
    pvinv = np.stack([c.pvinv for c in cams]).astype(np.float32)
    d = L.FitDesc()
    d.width, d.height, d.views = scene.W, scene.H, len(cams)
    d.V, d.T = scene.V, scene.T
    d.texwidth, d.texheight, d.channel, d.miplevel, d.sgnum = scene.tw, scene.th, scene.C, scene.L, scene.K
    d.pos, d.normals, d.uv = h(scene.pos), h(scene.normals), h(scene.uv)
    d.idx = h(scene.idx, np.uint32)
    d.texture = h(scene.tex)
    d.axis, d.sharpness, d.amplitude = h(scene.axis), h(scene.sharp), h(scene.amp)
    d.mats, d.pvinv, d.rot = h(mats), h(pvinv), h(np.eye(4, dtype=np.float32).reshape(16))
    d.target = h(target) if target is not None else None
    d.roughness, d.ior = scene.rough, scene.ior
    d.lr, d.beta1, d.beta2, d.eps = lr, betas[0], betas[1], eps
    d.fold_quarter = 1 if fold_quarter else 0
    d.op_by_op = 1 if op_by_op else 0
    return d, keep


# ---------------------------------------------------------------- tuning (vcl_set_tuning)
def set_tuning(raster_tile=16, bin_chunk=256, splat_rows=16):
    t = L.Tuning(raster_tile, bin_chunk, splat_rows)
    L.check(_lib().vcl_set_tuning(C.byref(t)), "set_tuning")


def get_tuning():
    t = L.Tuning()
    L.check(_lib().vcl_get_tuning(C.byref(t)), "get_tuning")
    return {"raster_tile": t.raster_tile, "bin_chunk": t.bin_chunk, "splat_rows": t.splat_rows}


class Fit:
    """One device-resident fit request (vcl_fit_* of include/vcl.h)."""

    def __init__(self, scene, target=None, lr=1e-2, betas=(0.9, 0.999), eps=1e-8, fold_quarter=True, op_by_op=False, views=None):
        cams = scene.cams if views is None else views
        d, self._keep = fit_desc(scene, target, lr, betas, eps, fold_quarter, op_by_op, cams)
        self.desc = d
        self.scene = scene
        self.views = len(cams)
        h_ = C.c_void_p()
        L.check(_lib().vcl_fit_create(C.byref(d), C.byref(h_)), "fit_create")
        self.h = h_
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            _lib().vcl_fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_target_from_render(self):
        L.check(_lib().vcl_fit_set_target_from_render(self.h, _s()), "fit_set_target")

    def set_tuning(self, raster_tile=16, bin_chunk=256, splat_rows=16):
        t = L.Tuning(raster_tile, bin_chunk, splat_rows)
        L.check(_lib().vcl_fit_set_tuning(self.h, C.byref(t)), "fit_set_tuning")

    def tuning(self):
        t = L.Tuning()
        L.check(_lib().vcl_fit_get_tuning(self.h, C.byref(t)), "fit_get_tuning")
        return {"raster_tile": t.raster_tile, "bin_chunk": t.bin_chunk, "splat_rows": t.splat_rows}

    def autotune(self, steps=3):
        """Pick the fastest kernel schedule for this request (state restored afterwards)."""
        t = L.Tuning()
        ms = C.c_double()
        L.check(_lib().vcl_fit_autotune(self.h, int(steps), C.byref(t), C.byref(ms)), "fit_autotune")
        return {"raster_tile": t.raster_tile, "bin_chunk": t.bin_chunk, "splat_rows": t.splat_rows, "ms_per_step": ms.value}

    def step(self, graph=False):
        if graph:
            L.check(_lib().vcl_fit_step_graph(self.h, _s()), "fit_step_graph")
        else:
            L.check(_lib().vcl_fit_step(self.h, _s()), "fit_step")

    def ptr(self, name):
        p = _lib().vcl_fit_buffer(self.h, name.encode())
        if not p:
            raise KeyError(name)
        return p

    def read(self, name, shape, dtype=np.float32):
        """Synchronous copy of a device buffer into numpy (tests and reporting only)."""
        out = np.empty(shape, dtype)
        L.check(_lib().vcl_fit_read(self.h, name.encode(), out.ctypes.data, out.nbytes, _s()), f"fit_read {name}")
        return out

    def write(self, name, arr):
        arr = np.ascontiguousarray(arr)
        L.check(_lib().vcl_fit_write(self.h, name.encode(), arr.ctypes.data, arr.nbytes, _s()), f"fit_write {name}")

    def profile_begin(self, max_launches=4096, only=None):
        L.check(_lib().vcl_fit_profile_begin(self.h, max_launches, only.encode() if only else None), "profile_begin")

    def profile_end(self):
        """{op name: (total ms, launches)} for the steps since profile_begin (synchronises)."""
        n = C.c_int()
        names = C.create_string_buffer(64 * 64)
        tot = (C.c_double * 64)()
        cnt = (C.c_int * 64)()
        L.check(_lib().vcl_fit_profile_end(self.h, 64, C.byref(n), names, 64, tot, cnt), "profile_end")
        out = {}
        for i in range(n.value):
            nm = names.raw[i * 64:(i + 1) * 64].split(b"\0")[0].decode()
            out[nm] = (tot[i], cnt[i])
        return out

    def device_bytes(self):
        return int(_lib().vcl_fit_device_bytes(self.h))


def step_signature(fit):
    """Hash of the kernels and launch shapes of one step of `fit` (equal: they can be grouped)."""
    s = C.c_uint64()
    L.check(_lib().vcl_fit_step_signature(fit.h, C.byref(s)), "fit_step_signature")
    return s.value


class Group:
    """One step of several Fit requests in shared launches (vcl_group_* of include/vcl.h)."""

    def __init__(self, fits):
        self.fits = list(fits)
        arr = (C.c_void_p * len(self.fits))(*[f.h.value for f in self.fits])
        h_ = C.c_void_p()
        L.check(_lib().vcl_group_create(arr, len(self.fits), _s(), C.byref(h_)), "group_create")
        self.h = h_

    def step(self, graph=False):
        fn = _lib().vcl_group_step_graph if graph else _lib().vcl_group_step
        L.check(fn(self.h, _s()), "group_step")

    def launches(self):
        return int(_lib().vcl_group_launches(self.h))

    def close(self):
        if getattr(self, "h", None):
            torch.cuda.synchronize()
            _lib().vcl_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
