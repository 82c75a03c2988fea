"""GPU parity: each HIP op (through the C ABI) against the CPU oracle on the same seeded inputs.

Bar: triangle IDs / coverage bit-exact; values within 1e-5 relative (FP32); gradients within
1e-3 relative.  The FP64 oracle error of both FP32 paths is printed where it is informative.
"""
import numpy as np
import pytest

from helpers import (F32, assert_close, forward_inputs, grid_mesh, o_aa, o_aa_bwd, o_interp, o_interp_bwd, o_normal_axis, o_project,
                     o_raster, o_raster_bwd, o_sg, o_sg_bwd, o_spec, o_tex, o_tex_bwd, o_view_axis, rel_err)
from vcl_amd import scene as S

pytestmark = pytest.mark.gpu


def T(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def N(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def cfg1(orc):
    sc = S.config1_scene(1)
    d = forward_inputs(sc)
    return sc, d


def idx_t(sc):
    import torch
    return T(sc.idx.astype(np.int32)).view(torch.int32)


# ------------------------------------------------------------------ project / rasterize
def test_project_bitexact(gpu, cfg1):
    sc, d = cfg1
    clip = N(gpu.project(T(sc.pos), T(sc.cams[0].mat)))
    assert np.array_equal(clip, d["clip"]), "project must be bit-exact (no FP contraction on either side)"


def _check_raster(gpu, clip, idx, W, H, views=1, what=""):
    import torch
    ref, refdb = o_raster(clip, idx, W, H, views)
    r = gpu.Rasterizer(W, H, views, clip.shape[-2], idx.shape[0])
    out, db = r.forward(T(clip), T(idx.astype(np.int32)).view(torch.int32))
    assert r.status(T(clip), T(idx.astype(np.int32)).view(torch.int32), out, db) == 0
    out, db = N(out), N(db)
    ids_g, ids_r = out[..., 3], ref[..., 3]
    assert np.array_equal(ids_g, ids_r), f"{what}: triangle ids differ at {int((ids_g != ids_r).sum())} pixels"
    assert_close(out[..., :3], ref[..., :3], 1e-5, f"{what} (u,v,z)")
    assert_close(db, refdb, 1e-5, f"{what} DB")
    return out, db


def test_raster_config1(gpu, cfg1):
    sc, d = cfg1
    out, db = _check_raster(gpu, d["clip"], sc.idx, sc.W, sc.H, what="config1")
    assert (out[..., 3] > 0).sum() > 30000
    # error of the float path against the FP64 yardstick
    r64, db64 = o_raster(d["clip"].astype(np.float64), sc.idx, sc.W, sc.H, dt=np.float64)
    print("raster uv rel err vs fp64:", rel_err(out[..., :2], r64[..., :2]), "DB:", rel_err(db, db64))


@pytest.mark.parametrize("W,H", [(64, 64), (67, 45), (256, 128)])
def test_raster_grid_edges_on_centres(gpu, W, H):
    # vertices on a lattice that puts many edges exactly through pixel centres
    clip, idx = grid_mesh(8)
    _check_raster(gpu, clip, idx, W, H, what=f"grid {W}x{H}")


def test_raster_soup_ties_and_clipping(gpu):
    rng = np.random.default_rng(7)
    n = 3000
    clip = rng.uniform(-1.3, 1.3, (3 * n, 4)).astype(F32)
    clip[:, 3] = rng.uniform(0.5, 2.0, 3 * n)
    clip[:, 2] = rng.uniform(-1.5, 1.5, 3 * n) * clip[:, 3]
    # a block of coplanar duplicates (ties resolved by smaller id)
    clip[300:600] = clip[0:300]
    # some triangles crossing the camera plane (w < 0) and huge ones (guard band)
    clip[900:960, 3] = -clip[900:960, 3]
    clip[960:990, :2] *= 60.0
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    idx[100:200] = idx[0:100]
    _check_raster(gpu, clip, idx, 96, 80, what="soup")


def test_raster_occlusion_layers(gpu):
    # heavy overdraw of large perspective triangles (whole-tile covers at many depths, so the
    # fine raster's occlusion culling is active): near-ties 1e-7 apart in depth, triangles
    # crossing w = 0 and vertex w varying across each triangle (z not affine in the piece)
    rng = np.random.default_rng(23)
    n = 600
    ctr = rng.uniform(-0.6, 0.6, (n, 1, 2))
    xy = ctr + rng.uniform(-1.2, 1.2, (n, 3, 2))
    w = rng.uniform(0.6, 1.8, (n, 3, 1))
    z = rng.uniform(-0.95, 0.95, (n, 1, 1)) + rng.uniform(-0.05, 0.05, (n, 3, 1))
    clip = np.concatenate([xy * w, z * w, w], 2).astype(F32)
    clip[1::7] = clip[0::7][: len(clip[1::7])]
    clip[1::7, :, 2] += (1e-7 * clip[1::7, :, 3]).astype(F32)
    clip[3::29, 0, 3] = -clip[3::29, 0, 3]
    clip = clip.reshape(-1, 4)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    for W, H in ((128, 96), (75, 53)):
        out, _ = _check_raster(gpu, clip, idx, W, H, what=f"occlusion layers {W}x{H}")
        assert (out[..., 3] > 0).mean() > 0.9


def test_raster_crumpled_mesh(gpu):
    # a closed mesh whose vertices are scattered as by a diverged fit: long triangles folded over
    # each other many layers deep
    unit, faces = S.icosphere(4)
    rng = np.random.default_rng(29)
    pos = (unit * (1.0 + rng.normal(0, 0.25, (len(unit), 1)))).astype(F32)
    cam = S.make_camera((0.0, 0.0, 3.0))
    clip = o_project(pos, cam.mat)[0]
    _check_raster(gpu, clip, faces.astype(np.uint32), 200, 152, what="crumpled mesh")


def test_raster_deep_bins(gpu):
    # thousands of small overlapping triangles stacked over a few tiles: bins far deeper than
    # one fine-raster chunk (256 pieces), so the partial depth keys of several workgroups merge
    rng = np.random.default_rng(11)
    n = 6000
    ctr = rng.uniform(-0.5, 0.5, (n, 1, 2))
    xy = ctr + rng.uniform(-0.12, 0.12, (n, 3, 2))
    w = rng.uniform(0.8, 1.5, (n, 3, 1))
    z = rng.uniform(-0.9, 0.9, (n, 3, 1))
    clip = np.concatenate([xy * w, z * w, w], 2).reshape(-1, 4).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    out, _ = _check_raster(gpu, clip, idx, 80, 72, what="deep bins")
    assert (out[..., 3] > 0).sum() > 1500


def test_raster_bin_overflow_fallback(gpu):
    # bins forced far below what the soup needs: overflowing tiles scan every piece, same result
    from vcl_amd import _lib
    rng = np.random.default_rng(5)
    n = 1500
    clip = rng.uniform(-1.2, 1.2, (3 * n, 4)).astype(F32)
    clip[:, 3] = rng.uniform(0.5, 2.0, 3 * n)
    clip[:, 2] = rng.uniform(-1.0, 1.0, 3 * n) * clip[:, 3]
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    lib = _lib.load()
    lib.vcl_rasterize_set_bin_capacity(5000)
    try:
        _check_raster(gpu, clip, idx, 72, 56, what="bin overflow")
    finally:
        lib.vcl_rasterize_set_bin_capacity(0)


def test_raster_multiview_and_empty(gpu, cfg1):
    sc, _ = cfg1
    cams = S.multiview_cameras(3)
    clip = o_project(sc.pos, np.stack([c.mat for c in cams]), views=3)
    _check_raster(gpu, clip, sc.idx, 64, 64, views=3, what="3 views")
    _check_raster(gpu, clip[:1], np.zeros((0, 3), np.uint32), 40, 40, what="no triangles")


def test_raster_fullscreen_quad(gpu):
    clip = np.array([[-1, -1, 0, 1], [1, -1, 0, 1], [1, 1, 0, 1], [-1, 1, 0, 1]], F32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    out, _ = _check_raster(gpu, clip, idx, 33, 17, what="quad")
    assert (out[..., 3] > 0).all()


def test_raster_backward(gpu, cfg1):
    import torch
    sc, d = cfg1
    rng = np.random.default_rng(3)
    grast = rng.normal(size=(1, sc.H, sc.W, 2)).astype(F32)
    ref = o_raster_bwd(d["clip"], sc.idx, d["rast"], grast)
    r = gpu.Rasterizer(sc.W, sc.H, 1, sc.V, sc.T)
    g = torch.zeros(1, sc.V, 4, device="cuda")
    r.backward(T(d["clip"]), idx_t(sc), T(d["rast"]), T(grast), g)
    assert_close(N(g), ref, 1e-3, "rasterize backward", atol_frac=1e-4)


# ------------------------------------------------------------------ interpolate
def test_interpolate(gpu, cfg1):
    sc, d = cfg1
    out, da = gpu.interpolate(T(d["rast"]), T(d["db"]), T(sc.uv), idx_t(sc))
    assert_close(N(out), d["uvp"], 1e-5, "interpolate out")
    assert_close(N(da), d["uvDA"], 1e-5, "interpolate DA")
    rng = np.random.default_rng(4)
    gout = rng.normal(size=(1, sc.H, sc.W, 3)).astype(F32)
    attr = rng.normal(size=(sc.V, 3)).astype(F32)
    ga_ref, gr_ref = o_interp_bwd(d["rast"], attr, sc.idx, gout)
    import torch
    ga = torch.zeros(sc.V, 3, device="cuda")
    gr = torch.zeros(1, sc.H, sc.W, 2, device="cuda")
    gpu.interpolate_backward(T(d["rast"]), T(attr), idx_t(sc), T(gout), ga, gr)
    assert_close(N(ga), ga_ref, 1e-3, "interpolate grad attr", atol_frac=1e-5)
    assert_close(N(gr), gr_ref, 1e-5, "interpolate grad rast")


# ------------------------------------------------------------------ texture
@pytest.mark.parametrize("mode", [0, 1])
def test_texture(gpu, cfg1, mode):
    import torch
    sc, d = cfg1
    levels = gpu.level_views(T(sc.tex), sc.tw, sc.th, sc.C, sc.L)
    out = gpu.texture(T(d["rast"]), T(d["uvp"]), T(d["uvDA"]), levels)
    assert_close(N(out), d["diff"], 1e-5, "texture fwd")
    rng = np.random.default_rng(5)
    gout = rng.normal(size=(1, sc.H, sc.W, sc.C)).astype(F32)
    gt_ref, guv_ref, gda_ref = o_tex_bwd(d["rast"], d["uvp"], d["uvDA"], sc.tex, sc.tw, sc.th, sc.C, sc.L, gout, mode)
    gflat = torch.zeros(sc.tex.size, device="cuda")
    glv = gpu.level_views(gflat, sc.tw, sc.th, sc.C, sc.L)
    guv = torch.zeros(1, sc.H, sc.W, 2, device="cuda")
    gda = torch.zeros(1, sc.H, sc.W, 4, device="cuda")
    gpu.texture_backward(T(d["rast"]), T(d["uvp"]), T(d["uvDA"]), levels, T(gout), glv, guv, gda, da_mode=mode)
    assert_close(N(gflat), gt_ref, 1e-3, "texture grad texels", atol_frac=1e-5)
    assert_close(N(guv), guv_ref, 1e-3, "texture grad uv", atol_frac=1e-5)
    assert_close(N(gda), gda_ref, 1e-3, "texture grad uvDA", atol_frac=1e-5)


def _tex_bwd_case(case, sc, d):
    """Inputs that steer the texel-grad summation of a 16 x 16 region through each of its forms:
    a compact patch of texels (fixed-slot window), scattered texels over many levels (hash table),
    all-zero output grads, output grads spanning 8 decades in one region, and non-finite ones."""
    rng = np.random.default_rng(11)
    rast, uv, da = d["rast"], d["uvp"], d["uvDA"].reshape(d["rast"].shape[:3] + (4,))
    views, H, W = rast.shape[:3]
    gout = rng.normal(size=(views, H, W, sc.C)).astype(F32)
    if case == "scattered":
        H = W = 64
        rast = np.zeros((2, H, W, 4), F32)
        rast[..., 3] = 1.0
        uv = rng.uniform(-0.1, 1.1, (2, H, W, 2)).astype(F32)
        da = (rng.normal(size=(2, H, W, 4)) * 10.0 ** rng.uniform(-4, -1, (2, H, W, 1))).astype(F32)
        gout = rng.normal(size=(2, H, W, sc.C)).astype(F32)
    elif case == "magnified":
        H = W = 48
        rast = np.zeros((1, H, W, 4), F32)
        rast[..., 3] = 1.0
        yy, xx = np.mgrid[0:H, 0:W].astype(F32)
        uv = np.stack([0.3 + xx * 1e-3, 0.6 + yy * 7e-4], -1)[None].astype(F32)
        da = np.broadcast_to(np.array([1e-3, 0.0, 0.0, 7e-4], F32), (1, H, W, 4)).copy()
        gout = rng.normal(size=(1, H, W, sc.C)).astype(F32)
    elif case == "zero":
        gout[:, :, : W // 2] = 0.0
    elif case == "range":
        # a smooth uv ramp at 2.5 level-0 texels per pixel (every pixel at level 1.32, away from
        # the level boundaries where float log2 of host and GPU may pick different levels), with
        # output grads spanning 8 decades within every region
        H = W = 64
        rast = np.zeros((1, H, W, 4), F32)
        rast[..., 3] = 1.0
        s = 2.5 / sc.tw
        yy, xx = np.mgrid[0:H, 0:W].astype(F32)
        uv = np.stack([0.2 + xx * s, 0.3 + yy * 0.8 * s], -1)[None].astype(F32)
        da = np.broadcast_to(np.array([s, 0.0, 0.0, 0.8 * s], F32), (1, H, W, 4)).copy()
        gout = (rng.normal(size=(1, H, W, sc.C)) * 10.0 ** rng.uniform(-8, 0, (1, H, W, 1))).astype(F32)
    elif case == "nonfinite":
        cov = np.argwhere(rast[..., 3] >= 1)
        v, y, x = cov[len(cov) // 3]
        gout[v, y, x, 1] = np.nan
        v, y, x = cov[2 * len(cov) // 3]
        gout[v, y, x, 2] = np.inf
    return rast, uv, da, gout


@pytest.mark.parametrize("case", ["scattered", "magnified", "zero", "range", "nonfinite"])
def test_texture_backward_summation_forms(gpu, cfg1, case):
    import torch
    sc, d = cfg1
    rast, uv, da, gout = _tex_bwd_case(case, sc, d)
    views, H, W = rast.shape[:3]
    levels = gpu.level_views(T(sc.tex), sc.tw, sc.th, sc.C, sc.L)
    gt_ref, guv_ref, gda_ref = o_tex_bwd(rast, uv, da, sc.tex, sc.tw, sc.th, sc.C, sc.L, gout, 0)
    gflat = torch.zeros(sc.tex.size, device="cuda")
    glv = gpu.level_views(gflat, sc.tw, sc.th, sc.C, sc.L)
    guv = torch.zeros(views, H, W, 2, device="cuda")
    gda = torch.zeros(views, H, W, 4, device="cuda")
    gpu.texture_backward(T(rast), T(uv), T(da), levels, T(gout), glv, guv, gda, da_mode=0)
    for name, got, ref in (("texels", N(gflat), gt_ref), ("uv", N(guv), guv_ref), ("uvDA", N(gda), gda_ref)):
        # non-finite entries in the same places, with the same signs of infinity
        assert np.array_equal(np.isnan(got), np.isnan(ref)), f"{case}: NaN pattern of grad {name}"
        assert np.array_equal(np.isposinf(got), np.isposinf(ref)) and np.array_equal(np.isneginf(got), np.isneginf(ref))
        fin = np.isfinite(ref)
        if case == "range" and name == "texels":
            # contributions spanning 8 decades within each region: every texel within 1e-3 of
            # its own size plus 1e-6 of the sum of its terms' sizes (the oracle's float sums are
            # good to about n 2^-24 of that) plus 1e-12 of the largest output grad (the GPU sums
            # a region in fixed point whose unit is 2^-51 of the region's largest output grad)
            mag = o_tex_bwd(rast, uv, da, sc.tex, sc.tw, sc.th, sc.C, sc.L, np.abs(gout), 0)[0]
            bad = np.abs(got - ref) > 1e-3 * np.abs(ref) + 1e-6 * mag + 1e-12 * np.abs(gout).max()
            assert not bad.any(), f"{case}: {int(bad.sum())} texel grads off, worst {np.max(np.abs(got - ref) / (mag + 1e-30))}"
            tiny = (mag > 0) & (mag < 1e-6 * mag.max())
            assert tiny.sum() > 5, "the case should hold texels whose terms are all tiny"
        else:
            assert_close(got[fin], ref[fin], 1e-3, f"{case}: grad {name}", atol_frac=1e-5)
    if case == "nonfinite":
        assert np.isnan(gt_ref).any()


def test_texture_nomip(gpu, cfg1):
    sc, d = cfg1
    lvl0 = sc.tex[: sc.tw * sc.th * sc.C]
    levels = gpu.level_views(T(lvl0), sc.tw, sc.th, sc.C, 1)
    out = gpu.texture(T(d["rast"]), T(d["uvp"]), T(d["uvDA"]), levels)
    ref = o_tex(d["rast"], d["uvp"], d["uvDA"], lvl0, sc.tw, sc.th, sc.C, 1)
    assert_close(N(out), ref, 1e-5, "texture no-mip fwd")
    # backward (TexturemapNoMipBackwardKernel, texturemap.cu:286-319): bilinear texel grads of
    # level 0 and the uv grad; no level, so no uvDA grad
    import torch
    rng = np.random.default_rng(6)
    gout = rng.normal(size=(1, sc.H, sc.W, sc.C)).astype(F32)
    gt_ref, guv_ref, _ = o_tex_bwd(d["rast"], d["uvp"], d["uvDA"], lvl0, sc.tw, sc.th, sc.C, 1, gout)
    gflat = torch.zeros(lvl0.size, device="cuda")
    glv = gpu.level_views(gflat, sc.tw, sc.th, sc.C, 1)
    guv = torch.zeros(1, sc.H, sc.W, 2, device="cuda")
    gpu.texture_backward(T(d["rast"]), T(d["uvp"]), T(d["uvDA"]), levels, T(gout), glv, guv, None)
    assert_close(N(gflat), gt_ref, 1e-3, "texture no-mip grad texels", atol_frac=1e-5)
    assert_close(N(guv), guv_ref, 1e-3, "texture no-mip grad uv", atol_frac=1e-5)


def test_mip_build_fold_bitexact(gpu, orc):
    import torch
    rng = np.random.default_rng(6)
    tw, th, C = 64, 32, 4
    L = S.mip_levels(tw, th, 16)
    n = sum((tw >> l) * (th >> l) * C for l in range(L))
    tex = np.zeros(n, F32)
    tex[: tw * th * C] = rng.uniform(0, 1, tw * th * C)
    ref = tex.copy()
    orc.lib().orc_mip_build_f32(tw, th, C, L, orc.P(ref))
    t = T(tex)
    gpu.build_mip(gpu.level_views(t, tw, th, C, L))
    assert np.array_equal(N(t), ref)
    g = rng.normal(size=n).astype(F32)
    for q in (1, 0):
        gref = g.copy()
        orc.lib().orc_mip_fold_f32(tw, th, C, L, orc.P(gref), q)
        gt = T(g)
        gpu.grad_sumup(gpu.level_views(gt, tw, th, C, L), quarter=bool(q))
        assert np.array_equal(N(gt), gref)
    del torch


@pytest.mark.parametrize("tw,th,nonfinite", [(32, 32, False), (96, 64, True), (200, 66, False), (130, 34, True)])
def test_texture_adam_update_bitexact(gpu, orc, tw, th, nonfinite):
    # widths that leave partial 128-texel workgroups, 2 to 6 levels, and NaN / Inf grads in the
    # coarse levels (the fold's AddNaNcheck per channel, which the level-1 quad splits over lanes)
    import torch
    rng = np.random.default_rng(8 + tw)
    C = 4
    L = S.mip_levels(tw, th, 16)
    n = sum((tw >> l) * (th >> l) * C for l in range(L))
    n0 = tw * th * C
    tex = S.build_mips_np(rng.uniform(0, 1, (th, tw, C)).astype(F32), L)
    g = rng.normal(size=n).astype(F32) * 0.1
    if nonfinite:
        n1 = (tw >> 1) * (th >> 1) * C
        g[n0 + 5] = np.nan
        g[n0 + n1 // 2 + 2] = np.inf
        if L > 2:
            g[n0 + n1 + 3] = -np.inf
        g[7] = np.nan
    m = rng.normal(size=n0).astype(F32) * 0.01
    v = rng.uniform(0, 0.01, n0).astype(F32)
    hyper = (0.9, 0.999, 0.9 ** 3, 0.999 ** 3, 1e-2, 1e-8)
    # oracle: fold -> adam(L0) -> clear -> clamp -> build
    rt, rg, rm, rv = tex.copy(), g.copy(), m.copy(), v.copy()
    lib = orc.lib()
    lib.orc_mip_fold_f32(tw, th, C, L, orc.P(rg), 1)
    lib.orc_adam_f32(n0, orc.P(rt), orc.P(rg), orc.P(rm), orc.P(rv), *hyper)
    lib.orc_clamp_f32(n0, orc.P(rt), 0.0, 1.0)
    lib.orc_mip_build_f32(tw, th, C, L, orc.P(rt))
    tt, gt, mt, vt = T(tex), T(g), T(m), T(v)
    gpu.texture_adam_update(gpu.level_views(tt, tw, th, C, L), gpu.level_views(gt, tw, th, C, L), mt, vt, hyper)
    assert np.array_equal(N(tt), rt)
    assert np.array_equal(N(mt), rm) and np.array_equal(N(vt), rv)
    assert not N(gt).any(), "texture grads must be cleared"
    del torch


def test_adam_multi_bitexact(gpu, orc):
    rng = np.random.default_rng(9)
    sizes = [1, 7, 3000, 100]
    ps = [rng.normal(size=s).astype(F32) for s in sizes]
    gs = [rng.normal(size=s).astype(F32) for s in sizes]
    ms = [rng.normal(size=s).astype(F32) * 0.1 for s in sizes]
    vs = [rng.uniform(0, 0.1, s).astype(F32) for s in sizes]
    gs[2][5] = np.nan
    hyper = (0.9, 0.999, 0.81, 0.998001, 1e-2, 1e-8)
    refs = []
    for p, g, m, v in zip(ps, gs, ms, vs):
        p, m, v = p.copy(), m.copy(), v.copy()
        orc.lib().orc_adam_f32(p.size, orc.P(p), orc.P(g), orc.P(m), orc.P(v), *hyper)
        refs.append((p, m, v))
    tp, tg, tm, tv = [T(x) for x in ps], [T(x) for x in gs], [T(x) for x in ms], [T(x) for x in vs]
    gpu.adam_step(tp, tg, tm, tv, hyper)
    for i in range(len(sizes)):
        np.testing.assert_array_equal(N(tp[i]), refs[i][0])
        np.testing.assert_array_equal(N(tm[i]), refs[i][1])
        np.testing.assert_array_equal(N(tv[i]), refs[i][2])


# ------------------------------------------------------------------ material
def test_normal_view_axis(gpu, cfg1):
    sc, d = cfg1
    n = gpu.normal_axis(T(d["rast"]), T(sc.normals), idx_t(sc), T(np.eye(4, dtype=F32).reshape(16)))
    assert_close(N(n), d["nrm"], 1e-5, "normal axis")
    cam = sc.cams[0]
    v = gpu.view_axis(sc.W, sc.H, cam.view.reshape(1, 16), cam.proj.reshape(1, 16))
    assert_close(N(v), d["view"], 1e-5, "view axis")


def test_sg_specular(gpu, cfg1):
    import torch
    sc, d = cfg1
    ax, sh, am = gpu.sg_specular(T(d["rast"]), T(d["nrm"]), T(d["view"]), T(d["rough"]), sc.ior)
    assert_close(N(ax), d["sAxis"], 1e-5, "spec axis")
    assert_close(N(sh), d["sSharp"], 1e-5, "spec sharpness")
    assert_close(N(am), d["sAmp"], 1e-5, "spec amplitude")
    rng = np.random.default_rng(10)
    ga = rng.normal(size=(1, sc.H, sc.W)).astype(F32)
    gs = rng.normal(size=(1, sc.H, sc.W)).astype(F32) * 1e-4
    ref = np.zeros((1, sc.H, sc.W), F32)
    import oracle as O
    O.lib().orc_sg_spec_bwd_f32(sc.W, sc.H, 1, O.P(d["rast"]), O.P(d["nrm"]), O.P(d["view"]), O.P(d["rough"]), sc.ior, O.P(ga), O.P(gs),
                                O.P(ref))
    gr = torch.zeros(1, sc.H, sc.W, device="cuda")
    gpu.sg_specular_backward(T(d["rast"]), T(d["nrm"]), T(d["view"]), T(d["rough"]), sc.ior, T(ga), T(gs), gr)
    assert_close(N(gr), ref, 1e-3, "spec roughness grad", atol_frac=1e-5)


def test_spherical_gaussian(gpu, cfg1):
    import torch
    sc, d = cfg1
    args = [T(d["rast"]), T(sc.axis), T(sc.sharp), T(sc.amp), T(d["nrm"]), T(d["diff"]), T(d["sAxis"]), T(d["sSharp"]), T(d["sAmp"])]
    out, de, se = gpu.spherical_gaussian(*args)
    assert_close(N(out), d["color"], 1e-5, "sg out")
    assert_close(N(de), d["denv"], 1e-5, "sg diffenv")
    assert_close(N(se), d["senv"], 1e-5, "sg specenv")
    rng = np.random.default_rng(11)
    gout = rng.normal(size=d["color"].shape).astype(F32)
    ref = o_sg_bwd(d["rast"], sc.axis, sc.sharp, sc.amp, d["nrm"], d["diff"], d["sAxis"], d["sSharp"], d["sAmp"], gout, spec_grads=True)
    K, C = sc.K, sc.C
    gd = torch.zeros_like(args[5])
    ga, gs, gm = torch.zeros(K, 3, device="cuda"), torch.zeros(K, device="cuda"), torch.zeros(K, C, device="cuda")
    gn = torch.zeros(1, sc.H, sc.W, 3, device="cuda")
    gsa = torch.zeros(1, sc.H, sc.W, 3, device="cuda")
    gss, gsm = torch.zeros(1, sc.H, sc.W, device="cuda"), torch.zeros(1, sc.H, sc.W, device="cuda")
    gpu.spherical_gaussian_backward(*args, T(gout), gd, ga, gs, gm, gn, gsa, gss, gsm)
    names = ["diffuse", "axis", "sharpness", "amplitude", "normal", "specAxis", "specSharpness", "specAmplitude"]
    for name, g, r in zip(names, [gd, ga, gs, gm, gn, gsa, gss, gsm], ref):
        assert_close(N(g), r, 1e-3, f"sg grad {name}", atol_frac=1e-4)
    # lobe-parallel path (no per-pixel grads requested): same grads, deterministic across runs
    res = []
    for _ in range(2):
        gd2 = torch.zeros_like(args[5])
        ga2, gs2, gm2 = torch.zeros_like(ga), torch.zeros_like(gs), torch.zeros_like(gm)
        gpu.spherical_gaussian_backward(*args, T(gout), gd2, ga2, gs2, gm2, denv=de)
        res.append([N(x) for x in (gd2, ga2, gs2, gm2)])
    for name, g2, r in zip(["diffuse", "axis", "sharpness", "amplitude"], res[0], ref[:4]):
        assert_close(g2, r, 1e-3, f"sg lobe-path grad {name}", atol_frac=1e-4)
    for a_, b_ in zip(res[0], res[1]):
        assert np.array_equal(a_, b_), "lobe grads must be deterministic"


def test_spherical_gaussian_adversarial(gpu, cfg1):
    # non-unit normals and lobe axes, very sharp specular lobes at some pixels, and NaN / Inf
    # amplitudes (AddNaNcheck drops those terms per channel): the fast C = 4 forward must still
    # match the oracle.  Lobe axes up to 2x unit length put the workgroups in the form that tests
    # every term; with unit axes, normals up to 3x unit length and spec amplitudes up to 1e6 they
    # stay in the bounded form (no per-term test, constant factors applied to the sums)
    sc, d = cfg1
    rng = np.random.default_rng(17)
    nrm = d["nrm"] * rng.uniform(0.3, 3.0, d["nrm"].shape[:-1] + (1,)).astype(F32)
    ss = d["sSharp"].copy()
    ss[..., ::7] = 1.0e5
    for amp_case in ("finite", "nonfinite", "bounded"):
        amp = sc.amp.copy()
        axis = (sc.axis * rng.uniform(0.2, 2.0, (sc.K, 1))).astype(F32) if amp_case != "bounded" else sc.axis
        sa = d["sAmp"].copy()
        if amp_case == "nonfinite":
            amp[3, 1] = np.nan
            amp[5, 2] = np.inf
        if amp_case == "bounded":
            amp = (amp * np.float32(3.0e4)).astype(F32)
            sa[..., ::5] = 1.0e6
        ref = o_sg(d["rast"], axis, sc.sharp, amp, nrm, d["diff"], d["sAxis"], ss, sa)
        out, de, se = gpu.spherical_gaussian(T(d["rast"]), T(axis), T(sc.sharp), T(amp), T(nrm), T(d["diff"]), T(d["sAxis"]), T(ss),
                                             T(sa))
        for name, g, r in zip(("out", "diffenv", "specenv"), (out, de, se), ref):
            g = N(g)
            assert np.array_equal(np.isfinite(g), np.isfinite(r)), f"sg {name} ({amp_case}): finiteness differs"
            assert_close(np.nan_to_num(g), np.nan_to_num(r), 1e-5, f"sg {name} ({amp_case})")


@pytest.mark.parametrize("K", [2, 3, 100, 200])
def test_spherical_gaussian_lobe_waves(gpu, cfg1, K):
    # more lobes than one wave of lobe lanes holds (two and four waves), sharpness over [0.5, 60]
    # with non-unit axes, very sharp specular lobes at some pixels and very wide ones at others:
    # the lobe-parallel backward stages the lobes by length and leaves out e^{-l-c} for a wave
    # and pixel chunk where it is below half an ulp; its lobe grads must match the oracle and
    # repeat exactly
    import torch
    sc, d = cfg1
    rng = np.random.default_rng(K)
    axis, sharp, amp = S.sg_lobes(K, sc.C, rng, 0.5, 60.0)
    axis = (axis * rng.uniform(0.5, 1.5, (K, 1))).astype(F32)
    ss = d["sSharp"].copy()
    ss[..., ::7] = 1.0e5
    ss[..., 3::11] = 0.5
    args = [T(d["rast"]), T(axis), T(sharp), T(amp), T(d["nrm"]), T(d["diff"]), T(d["sAxis"]), T(ss), T(d["sAmp"])]
    _, de, _ = gpu.spherical_gaussian(*args)
    gout = rng.normal(size=d["color"].shape).astype(F32)
    ref = o_sg_bwd(d["rast"], axis, sharp, amp, d["nrm"], d["diff"], d["sAxis"], ss, d["sAmp"], gout)
    res = []
    for _ in range(2):
        gd = torch.zeros_like(args[5])
        ga, gs, gm = torch.zeros(K, 3, device="cuda"), torch.zeros(K, device="cuda"), torch.zeros(K, sc.C, device="cuda")
        gpu.spherical_gaussian_backward(*args, T(gout), gd, ga, gs, gm, denv=de)
        res.append([N(x) for x in (gd, ga, gs, gm)])
    for name, g, r in zip(["diffuse", "axis", "sharpness", "amplitude"], res[0], ref[:4]):
        assert_close(g, r, 1e-3, f"sg lobe-wave grad {name} K={K}", atol_frac=1e-4)
    for a_, b_ in zip(res[0], res[1]):
        assert np.array_equal(a_, b_), "lobe grads must be deterministic"


# ------------------------------------------------------------------ antialias / loss
def test_antialias(gpu, cfg1):
    import torch
    sc, d = cfg1
    out = gpu.antialias(T(d["clip"]), idx_t(sc), T(d["rast"]), T(d["color"]))
    assert_close(N(out), d["image"], 1e-5, "antialias fwd")
    assert not np.array_equal(d["image"], d["color"]), "antialias must change edge pixels"
    rng = np.random.default_rng(12)
    gout = rng.normal(size=d["color"].shape).astype(F32)
    gin_ref, gclip_ref = o_aa_bwd(d["clip"], sc.idx, d["rast"], d["color"], gout)
    gclip = torch.zeros(1, sc.V, 4, device="cuda")
    gin = gpu.antialias_backward(T(d["clip"]), idx_t(sc), T(d["rast"]), T(d["color"]), T(gout), gclip)
    assert_close(N(gin), gin_ref, 1e-5, "antialias grad in")
    assert_close(N(gclip), gclip_ref, 1e-3, "antialias grad clip", atol_frac=1e-4)


@pytest.mark.parametrize("W,H,C", [(67, 45, 3), (33, 80, 1), (16, 16, 4)])
def test_antialias_ragged_views(gpu, W, H, C):
    # overlapping random triangles in three views, sizes that leave partial 16 x 16 tiles, so
    # pairs at tile edges, partial tiles and per-view offsets are all exercised
    import torch
    rng = np.random.default_rng(W * H + C)
    views, n = 3, 400
    ctr = rng.uniform(-0.9, 0.9, (views, n, 1, 2))
    xy = ctr + rng.uniform(-0.25, 0.25, (views, n, 3, 2))
    w = rng.uniform(0.8, 1.5, (views, n, 3, 1))
    z = rng.uniform(-0.9, 0.9, (views, n, 3, 1))
    clip = np.concatenate([xy * w, z * w, w], -1).reshape(views, 3 * n, 4).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    rast, _ = o_raster(clip, idx, W, H, views)
    color = rng.uniform(0, 1, (views, H, W, C)).astype(F32)
    idx_g = T(idx.astype(np.int32)).view(torch.int32)
    out = gpu.antialias(T(clip), idx_g, T(rast), T(color))
    ref = o_aa(clip, idx, rast, color)
    assert not np.array_equal(ref, color)
    assert_close(N(out), ref, 1e-5, f"antialias fwd {W}x{H}x{C}")
    gout = rng.normal(size=color.shape).astype(F32)
    gin_ref, gclip_ref = o_aa_bwd(clip, idx, rast, color, gout)
    gclip = torch.zeros(views, 3 * n, 4, device="cuda")
    gin = gpu.antialias_backward(T(clip), idx_g, T(rast), T(color), T(gout), gclip)
    assert_close(N(gin), gin_ref, 1e-5, f"antialias grad in {W}x{H}x{C}")
    assert_close(N(gclip), gclip_ref, 1e-3, f"antialias grad clip {W}x{H}x{C}", atol_frac=1e-4)


def test_mse(gpu):
    rng = np.random.default_rng(13)
    for n in (1, 1000, 256 * 256 * 4 + 3):
        a = rng.normal(size=n).astype(F32)
        b = rng.normal(size=n).astype(F32)
        g, loss = gpu.mse_loss(T(a), T(b))
        assert np.array_equal(N(g), a - b)
        ref = 0.5 * np.sum((a.astype(np.float64) - b) ** 2)
        assert abs(float(N(loss)[0]) - ref) <= 1e-5 * ref


# ------------------------------------------------------------------ whole step
def test_fit_step_config1(gpu, orc):
    """One fwd+bwd+Adam step of config 1 (icosphere L3, 256^2, 512^2x4 mip texture, 32 lobes)
    against the oracle's step: same loss, image, and updated parameters."""
    import ctypes as C
    sc = S.config1_scene(1)
    tgt_sc = S.config1_scene(2)
    # target from the second seed, rendered by the oracle forward
    target = forward_inputs(tgt_sc)["image"].reshape(sc.H, sc.W, sc.C).copy()
    fit = gpu.Fit(sc, target=target)
    fit.step()
    # oracle step
    L = orc.lib()
    s = orc.OrcScene()
    keep = {}

    def p(name, arr):
        keep[name] = np.ascontiguousarray(arr)
        return keep[name].ctypes.data_as(C.c_void_p)
    cam = sc.cams[0]
    s.W, s.H, s.V, s.T, s.C, s.K, s.tw, s.th, s.L = sc.W, sc.H, sc.V, sc.T, sc.C, sc.K, sc.tw, sc.th, sc.L
    s.pos = p("pos", sc.pos.copy()); s.normals = p("n", sc.normals); s.uv = p("uv", sc.uv); s.idx = p("idx", sc.idx)
    s.mat = p("mat", cam.mat); s.rot = p("rot", np.eye(4, dtype=F32).reshape(16)); s.pvinv = p("pv", cam.pvinv)
    s.tex = p("tex", sc.tex.copy()); s.axis = p("axis", sc.axis.copy()); s.sharp = p("sharp", sc.sharp.copy()); s.amp = p("amp", sc.amp.copy())
    s.rough = p("rough", np.full(sc.W * sc.H, sc.rough, F32)); s.ior = sc.ior; s.target = p("tgt", target); s.fold_quarter = 1
    for nm, n in (("pos", sc.V * 3), ("tex", sc.tw * sc.th * sc.C), ("axis", sc.K * 3), ("sharp", sc.K), ("amp", sc.K * sc.C)):
        setattr(s, "m_" + nm, p("m_" + nm, np.zeros(n, F32)))
        setattr(s, "v_" + nm, p("v_" + nm, np.zeros(n, F32)))
    s.lr, s.b1, s.b2, s.eps, s.b1t, s.b2t = 1e-2, 0.9, 0.999, 1e-8, 0.9, 0.999
    img = np.zeros((sc.H, sc.W, sc.C), F32)
    s.out_image = p("img", img)
    w = L.orc_work_create()
    L.orc_fit_step_f32(C.byref(s), w)
    L.orc_work_destroy(w)
    loss_g = float(fit.read("loss", (1,))[0])
    assert abs(loss_g - s.loss) <= 1e-5 * abs(s.loss), (loss_g, s.loss)
    assert_close(fit.read("image", img.shape), img, 1e-5, "step image")
    # first-step Adam moments hold m = 0.1*g: the gradients of every parameter
    for nm, n in (("pos", sc.V * 3), ("axis", sc.K * 3), ("sharp", sc.K), ("amp", sc.K * sc.C), ("tex", sc.tw * sc.th * sc.C)):
        assert_close(fit.read("m_" + nm, (n,)), keep["m_" + nm], 1e-3, f"step grad {nm}", atol_frac=1e-4)
    # updated parameters: a first Adam step moves each one by about lr*sign(g)
    for nm, key, base in (("pos", "pos", sc.pos), ("axis", "axis", sc.axis), ("sharpness", "sharp", sc.sharp),
                          ("amplitude", "amp", sc.amp), ("texture", "tex", sc.tex)):
        got = fit.read(nm, base.shape)
        ref = keep[key].reshape(base.shape)
        agree = np.mean(np.isclose(got, ref, rtol=1e-5, atol=1e-6))
        assert agree > 0.999, f"{nm}: only {agree:.5f} of the updated values agree"
    fit.close()


def test_fit_step_op_by_op_matches_fused(gpu, orc):
    """vcl_fit_desc.op_by_op = 1 runs the reference's separate launches (NormalAxis, ViewAxis,
    SGSpecular, Interpolate, Texturemap, their backwards in turn) instead of the fused passes;
    the per-pixel code is shared, so the first forward is identical and the gradients agree to
    the float-atomic summation order (which also makes the second step's parameters, hence its
    image, differ in the last bits)."""
    sc = S.config1_scene(1, W=128)
    target = forward_inputs(S.config1_scene(2, W=128))["image"].reshape(sc.H, sc.W, sc.C).copy()
    out = []
    for op_by_op in (False, True):
        fit = gpu.Fit(sc, target=target, op_by_op=op_by_op)
        fit.step()
        r = {"loss": float(fit.read("loss", (1,))[0]), "image": fit.read("image", (sc.H, sc.W, sc.C)),
             "m_pos": fit.read("m_pos", (sc.V * 3,)), "m_tex": fit.read("m_tex", (sc.tw * sc.th * sc.C,))}
        fit.step()
        r["loss2"] = float(fit.read("loss", (1,))[0])
        r["image2"] = fit.read("image", (sc.H, sc.W, sc.C))
        out.append(r)
        fit.close()
    f, o = out
    assert_close(o["image"], f["image"], 1e-6, "op-by-op image (step 1)")
    assert abs(o["loss"] - f["loss"]) <= 1e-6 * abs(f["loss"])
    # after step 1 the first moments are (1 - beta1) * gradient
    assert_close(o["m_pos"], f["m_pos"], 1e-3, "op-by-op position moments", atol_frac=1e-4)
    assert_close(o["m_tex"], f["m_tex"], 1e-3, "op-by-op texture moments", atol_frac=1e-4)
    # normwise: a silhouette pixel may move with the last-bit parameter differences
    assert rel_err(o["image2"], f["image2"]) < 1e-4
    assert abs(o["loss2"] - f["loss2"]) <= 1e-4 * abs(f["loss2"])
