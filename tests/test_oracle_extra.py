"""CPU pins of the oracle's second unit (oracle/vcl_oracle_extra.cpp): the ops beyond the fit
chain.  Each check takes an independent route -- a library routine (scipy.ndimage box sums
and correlations, numpy statistics), a textbook recurrence, a closed form on a symmetric input,
or central finite differences of the forward -- rather than retyping the oracle's formula."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from vcl_amd import scene as S

F32 = np.float32
D64 = np.float64
C1, C2 = float(F32(1e-4)), float(F32(1e-3))  # SSIM constants, float in the reference (loss.cu:153)


@pytest.fixture(scope="module", autouse=True)
def _lib(orc):
    O.set_threads(2)
    yield


def L():
    return O.lib()


def P(a):
    return O.P(a)


def uniform(idx, b, c):
    idx = np.ascontiguousarray(idx, np.uint32)
    out = np.zeros(idx.size, F32)
    L().orc_uniform(idx.size, P(idx), b, c, P(out))
    return out


# ------------------------------------------------------------------ random numbers, noise
def test_uniform_distribution():
    n = 1 << 20
    i = np.arange(n, dtype=np.uint32)
    u = uniform(i, 12345, 0xcafef00d)
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 2e-3
    assert abs(u.var() - 1.0 / 12) < 1e-3
    counts = np.histogram(u, bins=64, range=(0, 1))[0]
    chi2 = np.sum((counts - n / 64) ** 2 / (n / 64))
    assert chi2 < 120, chi2  # 63 degrees of freedom: p ~ 1e-5 at 120
    # one round of the mix leaves neighbouring inputs slightly correlated (consecutive indices
    # -0.016, consecutive seeds 0.014 over 2^20 samples, standard error 1e-3): a property of
    # the reference's generator, bounded here
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.03
    v = uniform(i, 12346, 0xcafef00d)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.03
    w = uniform(i, 12345, 0xdeadbeef)
    assert abs(np.corrcoef(u, w)[0, 1]) < 0.03


def test_noise_closed_forms():
    W, H, D, C_ = 19, 7, 2, 3
    rng = np.random.default_rng(0)
    x = rng.uniform(size=(D, H, W, C_)).astype(F32)
    out = np.zeros_like(x)
    L().orc_noise(W, H, D, C_, 0.0, P(x), P(out), 77)
    assert np.array_equal(out, x)
    L().orc_noise(W, H, D, C_, 1.0, P(x), P(out), 77)
    pidx = np.arange(W * H * D, dtype=np.uint32)
    for c in range(C_):
        np.testing.assert_allclose(out.reshape(-1, C_)[:, c], uniform(pidx, 77 + c, 0xcafef00d), rtol=0, atol=1e-7)
    L().orc_noise(W, H, D, C_, 0.25, P(x), P(out), 77)
    ref = x.reshape(-1, C_) + 0.25 * (np.stack([uniform(pidx, 77 + c, 0xcafef00d) for c in range(C_)], 1) - x.reshape(-1, C_))
    np.testing.assert_allclose(out.reshape(-1, C_), ref, atol=1e-6)


# ------------------------------------------------------------------ optimizers
def test_sgd_nadam_recurrences():
    rng = np.random.default_rng(1)
    n = 64
    p = rng.normal(size=n).astype(F32)
    g = rng.normal(size=n).astype(F32)
    p[3], g[5], g[6] = np.nan, np.inf, 1e38
    q = p.copy()
    L().orc_sgd(n, P(q), P(g), 0.1)
    want = p.astype(D64) - 0.1 * g.astype(D64)
    want[3] = 0.0                  # non-finite parameter reset, update dropped
    want[5] = p[5]                 # non-finite update dropped
    np.testing.assert_allclose(q, want.astype(F32), rtol=1e-6)
    # Nadam (Dozat 2016, the Keras schedule mu_t = mu (1 - 0.5 * 0.96^(0.004 t))): at kernel call
    # s the reference's host state is that of t = s + 1 (its setHyperParams starts one ahead)
    mu, rho, alpha, eps = 0.9, 0.999, 2e-3, 1e-8
    sched = lambda t: mu * (1 - 0.5 * 0.96 ** (0.004 * t))
    prod = lambda t: np.prod([sched(i) for i in range(1, t + 1)])
    param = rng.normal(size=n).astype(F32)
    m = np.zeros(n, F32)
    v = np.zeros(n, F32)
    pm, pv, pp = np.zeros(n), np.zeros(n), param.astype(D64)
    for s in range(1, 6):
        t = s + 1
        g = rng.normal(size=n).astype(F32)
        L().orc_nadam(n, P(param), P(g), P(m), P(v), mu, sched(t), sched(t + 1), rho, alpha, prod(t), prod(t + 1), rho ** t, eps)
        gd = g.astype(D64)
        pm = mu * pm + (1 - mu) * gd
        pv = rho * pv + (1 - rho) * gd * gd
        mhat = sched(t + 1) * pm / (1 - prod(t + 1)) + (1 - sched(t)) * gd / (1 - prod(t))
        vhat = pv / (1 - rho ** t)
        pp = pp - alpha * mhat / (np.sqrt(vhat) + eps)
        np.testing.assert_allclose(param, pp, rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------------ buffer utilities
def test_buffer_utilities():
    rng = np.random.default_rng(2)
    n = 1000
    x = rng.normal(size=n).astype(F32)
    y = np.zeros(n, F32)
    L().orc_linear(n, P(y), P(x), 2.0, -1.0)
    np.testing.assert_array_equal(y, x * F32(2) + F32(-1))
    L().orc_clamp_buf(n, P(y), P(x), -0.5, 0.5)
    np.testing.assert_array_equal(y, np.clip(x, -0.5, 0.5))
    L().orc_step_buf(n, P(y), P(x), 0.1)
    np.testing.assert_array_equal(y, (x >= 0.1).astype(F32))
    y = x.copy()
    L().orc_add_random(n, P(y), -2.0, 3.0, 9)
    np.testing.assert_allclose(y - x, -2.0 + 5.0 * uniform(np.arange(n), 9, 0xba5ec0de), atol=1e-5)
    L().orc_random_params(n, P(y), 4, 1.0, 100.0)
    np.testing.assert_allclose(y, 1.0 + 99.0 * uniform(np.arange(n), 4, 0xdeadbeef), rtol=1e-6)
    rows = rng.normal(size=(50, 4)).astype(F32)
    rows[7] = 0
    z = rows.copy()
    L().orc_normalize_rows(50, 4, P(z), P(z))
    nz = np.linalg.norm(rows, axis=1) > 0
    np.testing.assert_allclose(z[nz], rows[nz] / np.linalg.norm(rows[nz], axis=1, keepdims=True), rtol=1e-6)
    assert np.array_equal(z[7], rows[7])
    d = np.zeros((20000, 3), F32)
    L().orc_spherical_random(20000, P(d), 3)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, rtol=1e-5)
    assert np.all(np.abs(d.mean(0)) < 0.02)           # uniform on the sphere: zero mean
    assert abs(np.mean(d[:, 2] ** 2) - 1 / 3) < 0.01  # and E[z^2] = 1/3
    a, b = rng.normal(size=(30, 3)).astype(F32), rng.normal(size=(30, 3)).astype(F32)
    assert abs(L().orc_distance_error(30, 3, P(a), P(b)) - np.mean(np.linalg.norm(a.astype(D64) - b, axis=1))) < 1e-6


def test_bilinear_downsampling():
    """Level l = the level-(l-1) image correlated with [1 3 3 1]/8 along both axes (zero outside,
    i.e. dropped taps) and sampled at odd positions; channel 3 of a 4-channel texture untouched."""
    tw, th, C_ = 32, 16, 4
    Lv = 4
    rng = np.random.default_rng(3)
    tex = rng.uniform(size=sum((tw >> l) * (th >> l) * C_ for l in range(Lv))).astype(F32)
    before = tex.copy()
    L().orc_bilinear_downsampling(tw, th, C_, Lv, P(tex))
    w = np.array([1, 3, 3, 1]) / 8.0
    off = 0
    src = tex[: tw * th * C_].reshape(th, tw, C_).astype(D64)
    for l in range(1, Lv):
        off_prev = off
        off += (tw >> (l - 1)) * (th >> (l - 1)) * C_
        h, w_ = th >> l, tw >> l
        got = tex[off: off + h * w_ * C_].reshape(h, w_, C_)
        for c in range(3):
            r = ndimage.correlate1d(ndimage.correlate1d(src[..., c], w, axis=1, mode="constant"), w, axis=0, mode="constant")
            np.testing.assert_allclose(got[..., c], r[1::2, 1::2], rtol=1e-5, atol=1e-6)
        np.testing.assert_array_equal(got[..., 3], before[off: off + h * w_ * C_].reshape(h, w_, C_)[..., 3])
        src = got.astype(D64)


def test_sg_bake_normalize_disperse():
    W, H = 64, 32
    # one lobe along +z: the texel at px = W/2, py = H/2 looks along (0, 0, 1), so it holds the
    # amplitude; elsewhere amp * exp(lambda (cos phi cos theta - 1))
    axis = np.array([[0, 0, 1]], F32)
    sharp = np.array([7.0], F32)
    amp = np.array([[0.5, 1.0, 2.0]], F32)
    tex = np.zeros((H, W, 3), F32)
    L().orc_sg_bake(1, 3, P(axis), P(sharp), P(amp), W, H, P(tex))
    np.testing.assert_allclose(tex[H // 2, W // 2], amp[0], rtol=1e-6)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    phi = (1 - 2 * px / W) * math.pi
    theta = (2 * py / H - 1) * math.pi / 2
    ref = np.exp(7.0 * (np.cos(phi) * np.cos(theta) - 1))[..., None] * amp[0]
    np.testing.assert_allclose(tex, ref, rtol=2e-5, atol=1e-7)
    # normalize: unit axes, sharpness in [0, 8], amplitude in [0, 1]
    rng = np.random.default_rng(4)
    ax = rng.normal(size=(10, 3)).astype(F32)
    sh = rng.uniform(-5, 20, 10).astype(F32)
    am = rng.uniform(-1, 2, (10, 3)).astype(F32)
    a0, s0, m0 = ax.copy(), sh.copy(), am.copy()
    L().orc_sg_normalize(10, 3, P(ax), P(sh), P(am))
    np.testing.assert_allclose(ax, a0 / np.linalg.norm(a0, axis=1, keepdims=True), rtol=1e-6)
    np.testing.assert_array_equal(sh, np.clip(s0, 0, 8))
    np.testing.assert_array_equal(am, np.clip(m0, 0, 1))
    # disperse, two lobes: equal and opposite pushes along their separation, of size
    # 1 / max(|d|^2 * sharpness * |amplitude difference|^2, 1e-6)
    ax = np.array([[1, 0, 0], [0, 1, 0]], F32)
    sh = np.array([2.0, 2.0], F32)
    am = np.array([[0.2, 0.3, 0.4], [0.5, 0.1, 0.4]], F32)
    buf = np.zeros((2, 3), F32)
    a0 = ax.copy()
    L().orc_sg_disperse(2, 3, P(ax), P(sh), P(am), P(buf))
    d = a0[0] - a0[1]
    mag = 1.0 / max(np.dot(d, d) * 2.0 * np.sum((am[0] - am[1]) ** 2), 1e-6)
    np.testing.assert_allclose(buf[0], d / np.linalg.norm(d) * mag, rtol=1e-5)
    np.testing.assert_allclose(buf[1], -buf[0], rtol=1e-6)
    np.testing.assert_allclose(ax, a0 + buf, rtol=1e-6)


# ------------------------------------------------------------------ losses
def test_texture_loss():
    W, H, D = 16, 12, 3
    rng = np.random.default_rng(5)
    a = rng.normal(size=(H, W, D)).astype(F32)
    b = rng.normal(size=(H, W, D)).astype(F32)
    buf = np.zeros_like(a)
    loss = L().orc_texture_loss(W, H, D, P(a), P(b), P(buf))
    ys, xs = np.mgrid[0:H, 0:W]
    masked = ((xs < W // 4) | (xs >= W // 2)) & ((ys < W // 4) | (ys >= W // 2))
    ref = np.where(masked[..., None], 0.0, (a.astype(D64) - b) ** 2)
    np.testing.assert_allclose(buf, ref, rtol=1e-6)
    assert abs(loss - ref.sum() / ((W * H * D) // 2)) < 1e-9 * max(1, ref.sum())


def _ssim64(x, y, mode=0):
    H, W, D = x.shape
    g = np.zeros_like(x)
    buf = np.zeros_like(x)
    loss = L().orc_ssim_f64(W, H, D, P(x), P(y), P(g), P(buf), mode)
    return loss, g, buf


def test_ssim_forward_box_sums():
    """S from window sums taken by scipy (ones-kernel correlation, zero outside the image), with
    means and (co)variances over 1/25 as the reference defines them."""
    rng = np.random.default_rng(6)
    H, W, D = 17, 23, 3
    x = rng.uniform(size=(H, W, D))
    y = np.clip(x + 0.2 * rng.normal(size=x.shape), 0, 1)
    loss, _, buf = _ssim64(x, y)
    box = lambda a: ndimage.correlate(a, np.ones((11, 11)), mode="constant")
    for c in range(D):
        X, Y = x[..., c], y[..., c]
        n = box(np.ones_like(X))
        sx, sy, sxx, syy, sxy = box(X), box(Y), box(X * X), box(Y * Y), box(X * Y)
        mx, my = sx / 25, sy / 25
        vx = (sxx - 2 * mx * sx + n * mx * mx) / 25
        vy = (syy - 2 * my * sy + n * my * my) / 25
        cv = (sxy - my * sx - mx * sy + n * mx * my) / 25
        Sv = (2 * mx * my + C1) * (2 * cv + C2) / ((mx * mx + my * my + C1) * (vx + vy + C2))
        np.testing.assert_allclose(buf[..., c], 1 - Sv, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(loss, buf.sum() / (4 * x.size), rtol=1e-12)
    assert abs(_ssim64(x, x.copy())[0] - np.sum(1 - 1.0) / (4 * x.size)) < 1e-3  # identical images: S ~ 1


def test_ssim_ref_gradient_5x5():
    """On a 5 x 5 image every window is the whole image and 1/25 is its true mean, so
    S = N1 N2 / (D1 D2) with numpy's mean / var / covariance.  The reference's gradient
    (loss.cu:146-149) is then -(dN - (1 - S) dD) / (D1 D2) with dN, dD the derivatives of
    N1 N2 and D1 D2 w.r.t. the element, taken here by central differences."""
    rng = np.random.default_rng(7)
    x = rng.uniform(size=(5, 5, 2))
    y = rng.uniform(size=(5, 5, 2))
    _, g, buf = _ssim64(x, y, mode=0)
    for c in range(2):
        def parts(X):
            mx, my = X.mean(), y[..., c].mean()
            vx, vy = X.var(), y[..., c].var()
            cv = np.mean((X - mx) * (y[..., c] - my))
            return (2 * mx * my + C1) * (2 * cv + C2), (mx * mx + my * my + C1) * (vx + vy + C2)
        N0, D0 = parts(x[..., c])
        S0 = N0 / D0
        np.testing.assert_allclose(buf[..., c], 1 - S0, rtol=1e-10)
        h = 1e-6
        for i in range(5):
            for j in range(5):
                Xp, Xm = x[..., c].copy(), x[..., c].copy()
                Xp[i, j] += h
                Xm[i, j] -= h
                (Np, Dp), (Nm, Dm) = parts(Xp), parts(Xm)
                dN, dD = (Np - Nm) / (2 * h), (Dp - Dm) / (2 * h)
                want = -(dN - (1 - S0) * dD) / D0
                assert abs(g[i, j, c] - want) <= 1e-6 * max(1, abs(want)), (i, j, c, g[i, j, c], want)


def test_ssim_exact_gradient_fd():
    rng = np.random.default_rng(8)
    x = rng.uniform(size=(9, 13, 2))
    y = np.clip(x + 0.3 * rng.normal(size=x.shape), 0, 1)
    _, g, _ = _ssim64(x, y, mode=1)
    f = lambda xx: _ssim64(xx, y)[2].sum()
    h = 1e-6
    for (i, j, c) in [(0, 0, 0), (4, 6, 1), (8, 12, 0), (3, 1, 1), (6, 10, 0), (2, 7, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[i, j, c] += h
        xm[i, j, c] -= h
        fd = (f(xp) - f(xm)) / (2 * h)
        assert abs(g[i, j, c] - fd) <= 1e-5 * max(1, abs(fd)), (i, j, c, g[i, j, c], fd)


# ------------------------------------------------------------------ filter
def _weights(sigma):
    k = min(int(math.ceil(4 * sigma)), 127)
    i = np.arange(-k, k + 1)
    w = np.exp(-i * i / (2 * sigma * sigma))
    return w / w.sum()


@pytest.mark.parametrize("sigma", [0.7, 1.6, 3.3])
def test_filter_forward_scipy(sigma):
    rng = np.random.default_rng(9)
    D, H, W, C_ = 2, 30, 40, 3
    x = rng.normal(size=(D, H, W, C_))
    out = np.zeros_like(x)
    L().orc_filter_fwd_f64(W, H, D, C_, sigma, P(x), P(out))
    w = _weights(sigma)
    ref = ndimage.correlate1d(ndimage.correlate1d(x, w, axis=2, mode="constant"), w, axis=1, mode="constant")
    np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-12)


def test_filter_backward_adjoint_and_sigma_fd():
    rng = np.random.default_rng(10)
    D, H, W, C_ = 1, 21, 26, 2
    sigma = 1.3  # k = ceil(5.2) = 6 stays fixed within the finite-difference step
    x = rng.normal(size=(D, H, W, C_))
    gy = rng.normal(size=x.shape)
    gin = np.zeros_like(x)
    gs = L().orc_filter_bwd_f64(W, H, D, C_, sigma, P(x), P(gy), P(gin))
    fx = np.zeros_like(x)
    L().orc_filter_fwd_f64(W, H, D, C_, sigma, P(x), P(fx))
    z = rng.normal(size=x.shape)
    fz = np.zeros_like(x)
    L().orc_filter_fwd_f64(W, H, D, C_, sigma, P(z), P(fz))
    ginz = np.zeros_like(x)
    L().orc_filter_bwd_f64(W, H, D, C_, sigma, P(z), P(gy), P(ginz))
    assert abs(np.sum(fz * gy) - np.sum(z * ginz)) < 1e-10 * np.abs(fz * gy).sum()  # <F z, g> = <z, F^T g>

    def loss(s):
        o = np.zeros_like(x)
        L().orc_filter_fwd_f64(W, H, D, C_, s, P(x), P(o))
        return np.sum(o * gy)
    h = 1e-6
    fd = (loss(sigma + h) - loss(sigma - h)) / (2 * h)
    assert abs(gs - fd) <= 1e-6 * max(1, abs(fd)), (gs, fd)


# ------------------------------------------------------------------ compute normal
def _cn(pos, idx, mode):
    out = np.zeros_like(pos)
    L().orc_compute_normal_fwd_f64(pos.shape[0], P(pos), idx.shape[0], P(idx), mode, P(out))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_compute_normal_closed_forms(mode):
    # regular icosahedron (circumradius 1, edge 1.05, so |e1 . e2| < 1 for the raw-edge weight):
    # by symmetry every vertex normal is the radial direction
    v, f = S.icosphere(0)
    v = np.ascontiguousarray(v, D64)
    f = np.ascontiguousarray(f, np.uint32)
    n = _cn(v, f, mode)
    np.testing.assert_allclose(n, v / np.linalg.norm(v, axis=1, keepdims=True), atol=1e-12)
    # planar grid in z = 0, counter-clockwise seen from +z: every normal is +z
    from helpers import grid_mesh
    clip, idx = grid_mesh(5, jitter=0.05, seed=1)
    pos = np.ascontiguousarray(np.c_[clip[:, :2] * 0.3, np.zeros(len(clip))], D64)
    n = _cn(pos, idx, mode)
    np.testing.assert_allclose(n, np.tile([0, 0, 1.0], (len(pos), 1)), atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1])
def test_compute_normal_backward_fd(mode):
    v, f = S.icosphere(1)
    rng = np.random.default_rng(11)
    pos = np.ascontiguousarray(v * 0.6 + rng.normal(scale=0.02, size=v.shape), D64)
    f = np.ascontiguousarray(f, np.uint32)
    g = rng.normal(size=pos.shape)
    gp = np.zeros_like(pos)
    L().orc_compute_normal_bwd_f64(pos.shape[0], P(pos), f.shape[0], P(f), mode, P(g), P(gp))
    h = 1e-6
    for vi in (0, 5, 17, 33):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[vi, k] += h
            pm[vi, k] -= h
            fd = (np.sum(_cn(pp, f, mode) * g) - np.sum(_cn(pm, f, mode) * g)) / (2 * h)
            assert abs(gp[vi, k] - fd) <= 1e-5 * max(1, abs(fd)), (vi, k, gp[vi, k], fd)


# ------------------------------------------------------------------ debug raster
def test_debug_raster():
    from helpers import o_raster
    W = H = 64
    v = np.array([[-0.9, -0.8, 0.1, 1.0], [0.85, -0.9, 0.2, 1.0], [0.0, 0.9, 0.3, 1.0], [0.5, 0.5, -0.5, 1.0]], F32)
    idx = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    rast, db = o_raster(v, idx, W, H)
    P_ = W * H
    ids = np.zeros((P_, 4), F32)
    L().orc_raster_idhash(P_, P(rast), P(ids))
    r = rast.reshape(P_, 4)
    cov = r[:, 3] > 0
    assert np.all(ids[~cov] == 0) and np.all(ids[cov, 3] == 1)
    assert np.all((ids[cov, :3] >= 0) & (ids[cov, :3] <= 1))
    for t in (1, 2):  # one colour per triangle, different triangles differ
        sel = r[:, 3] == t
        if sel.any():
            assert np.all(ids[sel] == ids[sel][0])
    c1, c2 = ids[r[:, 3] == 1][0], ids[r[:, 3] == 2][0]
    assert np.abs(c1 - c2).max() > 1e-3
    wf = np.zeros((P_, 4), F32)
    L().orc_raster_wireframe(P_, P(rast), P(db), P(wf))
    assert np.all(wf[~cov] == 0) and np.all(wf[cov, 3] == 1)
    # triangle 0 has w = 1: its barycentrics are affine in the pixel, with the constant screen
    # gradients of the inverse of its 2 x 2 edge matrix (as in the affine raster pin); the GLSL
    # shading (rasterize.cpp:340-345) on those gives the expected colour
    A = np.array([[v[0, 0] - v[2, 0], v[1, 0] - v[2, 0]], [v[0, 1] - v[2, 1], v[1, 1] - v[2, 1]]], D64)
    Ai = np.linalg.inv(A)
    du = np.array([Ai[0, 0], Ai[0, 1]]) * 2 / W  # d(u)/d(pixel x, pixel y)
    dv = np.array([Ai[1, 0], Ai[1, 1]]) * 2 / H
    d = 0.5 * np.array([np.abs(du).sum(), np.abs(dv).sum(), np.abs(du + dv).sum()])
    sel = r[:, 3] == 1
    var = np.stack([r[sel, 0], r[sel, 1], 1 - r[sel, 0] - r[sel, 1]], 1).astype(D64)
    e = 1 - np.min(1 / (1 + np.exp(16 * (d[None, :] - var))), axis=1)
    np.testing.assert_allclose(wf[sel, 0], e, atol=2e-6)
    assert np.all(wf[sel, 0] == wf[sel, 1]) and np.all(wf[sel, 1] == wf[sel, 2])
    assert (var.min(1) < 0.002).sum() > 5 and np.all(e[var.min(1) < 0.002] > 0.5)  # lines on the edges
