"""GPU parity of the ops beyond the fit chain (vcl_amd/csrc/extra.hip) against the oracle
(oracle/vcl_oracle_extra.cpp) on the same seeded inputs.  Integer-mixed random numbers, the
elementwise utilities, SGD / Nadam (double arithmetic, no contraction) and the id hash are
bit-exact; ops with transcendental functions or float atomics are held to 1e-5 (values) and
1e-3 (gradients)."""
import numpy as np
import pytest

import oracle as O
from helpers import F32, assert_close, o_project, o_raster
from vcl_amd import scene as S

pytestmark = pytest.mark.gpu


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def N(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def L():
    return O.lib()


def P(a):
    return O.P(a)


def test_noise_and_random_utilities(gpu, orc):
    rng = np.random.default_rng(0)
    x = rng.uniform(size=(2, 33, 47, 3)).astype(F32)
    out = gpu.noise(T(x), 0.3, 1234)
    ref = np.zeros_like(x)
    L().orc_noise(47, 33, 2, 3, 0.3, P(x), P(ref), 1234)
    assert np.array_equal(N(out), ref)
    n = 100003
    y = rng.normal(size=n).astype(F32)
    t = T(y)
    gpu.buffer_add_random(t, -1.0, 2.0, 55)
    ry = y.copy()
    L().orc_add_random(n, P(ry), -1.0, 2.0, 55)
    assert np.array_equal(N(t), ry)
    gpu.random_params(t, 66, 1.0, 50.0)
    L().orc_random_params(n, P(ry), 66, 1.0, 50.0)
    assert np.array_equal(N(t), ry)
    d = gpu.zeros(5000, 3)
    gpu.buffer_spherical_random(d, 7)
    rd = np.zeros((5000, 3), F32)
    L().orc_spherical_random(5000, P(rd), 7)
    np.testing.assert_allclose(N(d), rd, atol=2e-6)


def test_elementwise_utilities(gpu, orc):
    rng = np.random.default_rng(1)
    n = 70001
    x = rng.normal(size=n).astype(F32)
    out = gpu.zeros(n)
    for fn, args, ofn, oargs in ((gpu.buffer_linear, (1.5, -0.25), L().orc_linear, (1.5, -0.25)),
                                 (gpu.buffer_clamp, (-0.3, 0.7), L().orc_clamp_buf, (-0.3, 0.7)),
                                 (gpu.buffer_step, (0.1,), L().orc_step_buf, (0.1,))):
        fn(out, T(x), *args)
        ref = np.zeros_like(x)
        ofn(n, P(ref), P(x), *oargs)
        assert np.array_equal(N(out), ref), fn.__name__
    rows = rng.normal(size=(999, 3)).astype(F32)
    rows[4] = 0
    t = T(rows)
    gpu.buffer_normalize(t, t, 3)
    ref = rows.copy()
    L().orc_normalize_rows(999, 3, P(ref), P(ref))
    assert np.array_equal(N(t), ref)
    t = gpu.zeros(300, 4)
    gpu.buffer_set_value(t, [0.75, 0.6, 0.55, 1.0])
    assert np.array_equal(N(t), np.tile(np.array([0.75, 0.6, 0.55, 1.0], F32), (300, 1)))
    a, b = rng.normal(size=(50000, 3)).astype(F32), rng.normal(size=(50000, 3)).astype(F32)
    de = float(N(gpu.distance_error(T(a), T(b)))[0])
    assert abs(de - L().orc_distance_error(50000, 3, P(a), P(b))) <= 1e-6 * de


def test_sgd_nadam_bitexact(gpu, orc):
    rng = np.random.default_rng(2)
    n = 40000
    p = rng.normal(size=n).astype(F32)
    g = rng.normal(size=n).astype(F32)
    p[10], g[11] = np.nan, np.inf
    tp = T(p)
    gpu.sgd_step(tp, T(g), 0.05)
    rp = p.copy()
    L().orc_sgd(n, P(rp), P(g), 0.05)
    assert np.array_equal(N(tp), rp, equal_nan=True)
    m, v = np.zeros(n, F32), np.zeros(n, F32)
    tp, tm, tv = T(p), T(m), T(v)
    rp = p.copy()
    args = (0.9, 0.89, 0.895, 0.999, 2e-3, 0.81, 0.72, 0.998, 1e-8)
    for _ in range(3):
        g = rng.normal(size=n).astype(F32)
        gpu.nadam_step(tp, T(g), tm, tv, *args)
        L().orc_nadam(n, P(rp), P(g), P(m), P(v), *args)
    assert np.array_equal(N(tp), rp, equal_nan=True)
    assert np.array_equal(N(tm), m) and np.array_equal(N(tv), v)


def test_texture_helpers(gpu, orc):
    import torch
    rng = np.random.default_rng(3)
    tw, th, C_, Lv = 64, 32, 4, 5
    sizes = [(tw >> l) * (th >> l) * C_ for l in range(Lv)]
    tex = rng.uniform(size=sum(sizes)).astype(F32)
    t = T(tex)
    lv = gpu.level_views(t, tw, th, C_, Lv)
    gpu.bilinear_downsampling(lv)
    ref = tex.copy()
    L().orc_bilinear_downsampling(tw, th, C_, Lv, P(ref))
    assert np.array_equal(N(t), ref)
    # SG bake into a 3-channel level stack (higher levels cleared)
    K = 40
    axis, sharp, amp = S.sg_lobes(K, 3, rng)
    sharp = (sharp * 2).astype(F32)
    bt, bh = 128, 64
    flat = torch.full((sum((bt >> l) * (bh >> l) * 3 for l in range(2)),), 5.0, device="cuda")
    lv = gpu.level_views(flat, bt, bh, 3, 2)
    gpu.sg_bake(T(axis), T(sharp), T(amp), lv)
    ref = np.zeros((bh, bt, 3), F32)
    L().orc_sg_bake(K, 3, P(axis), P(sharp), P(amp), bt, bh, P(ref))
    got = N(flat)
    assert_close(got[: bt * bh * 3].reshape(bh, bt, 3), ref, 1e-5, "sg bake")
    assert np.all(got[bt * bh * 3:] == 0)
    # SG normalize / disperse
    ax = rng.normal(size=(K, 3)).astype(F32)
    sh = rng.uniform(-2, 12, K).astype(F32)
    am = rng.uniform(-0.5, 1.5, (K, 3)).astype(F32)
    ta, ts, tm = T(ax), T(sh), T(am)
    gpu.sg_normalize(ta, ts, tm)
    L().orc_sg_normalize(K, 3, P(ax), P(sh), P(am))
    assert np.array_equal(N(ta), ax) and np.array_equal(N(ts), sh) and np.array_equal(N(tm), am)
    buf = gpu.sg_disperse(ta, ts, tm)
    rbuf = np.zeros((K, 3), F32)
    L().orc_sg_disperse(K, 3, P(ax), P(sh), P(am), P(rbuf))
    assert_close(N(buf), rbuf, 1e-6, "sg disperse force")
    assert_close(N(ta), ax, 1e-6, "sg disperse axes")


@pytest.mark.parametrize("shape", [(64, 96, 3), (33, 17, 1), (200, 256, 4)])
def test_ssim_and_texture_loss(gpu, orc, shape):
    H, W, D = shape
    rng = np.random.default_rng(4)
    x = rng.uniform(size=shape).astype(F32)
    y = np.clip(x + 0.2 * rng.normal(size=shape), 0, 1).astype(F32)
    for mode in (0, 1):
        g, loss, m = gpu.ssim_loss(T(x), T(y), mode)
        rg, rm = np.zeros_like(x), np.zeros_like(x)
        rl = L().orc_ssim_f32(W, H, D, P(x), P(y), P(rg), P(rm), mode)
        assert_close(N(m), rm, 1e-5, f"ssim map (mode {mode})")
        assert_close(N(g), rg, 1e-5 if mode == 0 else 1e-3, f"ssim grad (mode {mode})", atol_frac=1e-5)
        assert abs(float(N(loss)[0]) - rl) <= 1e-5 * abs(rl), (mode, float(N(loss)[0]), rl)
    loss, m = gpu.texture_loss(T(x), T(y))
    rm = np.zeros_like(x)
    rl = L().orc_texture_loss(W, H, D, P(x), P(y), P(rm))
    assert np.array_equal(N(m), rm)
    assert abs(float(N(loss)[0]) - rl) <= 1e-6 * abs(rl)


@pytest.mark.parametrize("sigma", [0.8, 2.5, 40.0])
def test_filter(gpu, orc, sigma):
    rng = np.random.default_rng(5)
    D, H, W, C_ = 2, 48, 80, 3
    x = rng.normal(size=(D, H, W, C_)).astype(F32)
    sd = T(np.array([sigma], F32))
    out = gpu.filter_forward(T(x), sd)
    ref = np.zeros_like(x)
    L().orc_filter_fwd_f32(W, H, D, C_, sigma, P(x), P(ref))
    assert_close(N(out), ref, 1e-5, f"filter fwd sigma {sigma}")
    gy = rng.normal(size=x.shape).astype(F32)
    gin, gs = gpu.filter_backward(T(x), sd, T(gy))
    rgin = np.zeros_like(x)
    rgs = L().orc_filter_bwd_f32(W, H, D, C_, sigma, P(x), P(gy), P(rgin))
    assert_close(N(gin), rgin, 1e-5, "filter grad in")
    r64 = L().orc_filter_bwd_f64(W, H, D, C_, sigma, P(x.astype(np.float64)), P(gy.astype(np.float64)), P(np.zeros(x.shape)))
    gsv = float(N(gs)[0])
    print(f"sigma {sigma}: d/dsigma GPU {gsv:.6e} float oracle {rgs:.6e} FP64 {r64:.6e}")
    assert abs(gsv - rgs) <= 1e-3 * abs(rgs) + 1e-6, (gsv, rgs)


@pytest.mark.parametrize("mode", [0, 1])
def test_compute_normal(gpu, orc, mode):
    v, f = S.icosphere(4)
    rng = np.random.default_rng(6)
    pos = (v * 0.7 + rng.normal(scale=0.01, size=v.shape)).astype(F32)
    f = np.ascontiguousarray(f, np.uint32)
    g = rng.normal(size=pos.shape).astype(F32)
    out, gpos = gpu.compute_normal(T(pos), T(f.view(np.int32)), mode, T(g))
    ref = np.zeros_like(pos)
    L().orc_compute_normal_fwd_f32(pos.shape[0], P(pos), f.shape[0], P(f), mode, P(ref))
    assert_close(N(out), ref, 1e-5, "compute normal")
    rg = np.zeros_like(pos)
    L().orc_compute_normal_bwd_f32(pos.shape[0], P(pos), f.shape[0], P(f), mode, P(g), P(rg))
    assert_close(N(gpos), rg, 1e-3, "compute normal grad", atol_frac=1e-5)


def test_debug_raster(gpu, orc):
    sc = S.config1_scene(1, W=128)
    clip = o_project(sc.pos, sc.cams[0].mat)[0]
    rast, db = o_raster(clip, sc.idx, 128, 128)
    ids = gpu.rasterize_idhash(T(rast))
    ref = np.zeros((128 * 128, 4), F32)
    L().orc_raster_idhash(128 * 128, P(rast), P(ref))
    assert np.array_equal(N(ids).reshape(-1, 4), ref)
    wf = gpu.rasterize_wireframe(T(rast), T(db))
    L().orc_raster_wireframe(128 * 128, P(rast), P(db), P(ref))
    assert_close(N(wf).reshape(-1, 4), ref, 1e-5, "wireframe")
