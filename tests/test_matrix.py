"""Rotation / Camera host math (include/vcl_matrix.hpp through the host-only C ABI; no GPU)
pinned by library routines, textbook forms and finite differences:
- getRotation (matrix.cu:43-53) equals scipy's quaternion -> matrix for unit quaternions;
- getView (matrix.cu:175-180) maps the eye to the origin and keeps R(q) as its rotation;
- setCam's quaternion (matrix.cu:182-192) reproduces the textbook look-at matrix (rows s, u, f
  with f = normalize(eye - center));
- the projection (matrix.cu:211-219) equals the textbook glFrustum of the symmetric frustum
  with half-height `size` and half-width `size * aspect` at the near plane;
- Camera::backward's exact mode and Rotation::backward's exact mode equal central differences;
  the reference modes are its stated products (matrix.cu:234-239) and half the derivative;
- Rotation::step is the FP64 Adam recurrence with the rhom^(t+1) bias correction, then
  renormalisation."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation as SR

from vcl_amd import _lib as L

F32 = np.float32


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libvcl_amd.so not built")
    return L.load()


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def m4(flat16):
    """column-major 16 floats -> the math matrix"""
    return np.asarray(flat16, np.float64).reshape(4, 4).T


def flat(M):
    return np.ascontiguousarray(np.asarray(M).T.reshape(16), F32)


def rot(lib, q):
    out = np.zeros(16, F32)
    assert lib.vcl_rotation_matrix(P(np.ascontiguousarray(q, F32)), P(out)) == 0
    return m4(out)


def unit_quats(n, seed):
    q = np.random.default_rng(seed).normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(F32)  # (x, y, z, w)


def test_rotation_matrix_matches_scipy(lib):
    for q in unit_quats(20, 1):
        R = rot(lib, q)
        ref = SR.from_quat(q.astype(np.float64)).as_matrix()  # scipy order (x, y, z, w), scalar last
        np.testing.assert_allclose(R[:3, :3], ref, atol=2e-6)
        np.testing.assert_allclose(R[3], [0, 0, 0, 1])
        np.testing.assert_allclose(R[:3, 3], 0)


def test_view_matrix(lib):
    rng = np.random.default_rng(2)
    for q in unit_quats(10, 3):
        eye = rng.normal(size=3).astype(F32)
        out = np.zeros(16, F32)
        assert lib.vcl_view_matrix(P(q), P(eye), P(out)) == 0
        V = m4(out)
        np.testing.assert_allclose(V[:3, :3], rot(lib, q)[:3, :3], atol=1e-7)
        np.testing.assert_allclose(V @ np.r_[eye, 1.0], [0, 0, 0, 1], atol=2e-6)


def _textbook_look_at(eye, center, up):
    f = eye - center
    f /= np.linalg.norm(f)
    s = np.cross(up, f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    M = np.eye(4)
    M[0, :3], M[1, :3], M[2, :3] = s, u, f
    M[:3, 3] = -M[:3, :3] @ eye
    return M


def test_look_at_quaternion_gives_textbook_view(lib):
    rng = np.random.default_rng(4)
    cases = [((0, 0, 3.5), (0, 0, -1)), ((0, 0, -0.35), (0, 0, 0)), ((0.35, 0, 0.1), (0, 0, 0))]
    cases += [(tuple(rng.normal(size=3) * 3), tuple(rng.normal(size=3) * 0.2)) for _ in range(30)]
    for eye, center in cases:
        eye = np.array(eye, F32)
        center = np.array(center, F32)
        up = np.array([0, 1, 0], F32)
        q = np.zeros(4, F32)
        assert lib.vcl_look_at_quat(P(eye), P(center), P(up), P(q)) == 0
        assert abs(np.linalg.norm(q) - 1) < 1e-5
        out = np.zeros(16, F32)
        assert lib.vcl_view_matrix(P(q), P(eye), P(out)) == 0
        ref = _textbook_look_at(eye.astype(np.float64), center.astype(np.float64), up.astype(np.float64))
        np.testing.assert_allclose(m4(out), ref, atol=5e-5 * max(1, np.abs(eye).max()))


def test_projection_is_gl_frustum(lib):
    for size, aspect, n, f in ((0.5, 1.0, 1.0, 10.0), (0.05, 1.0, 0.1, 10.0), (0.1 * np.tan(np.pi / 8), 1.5, 0.1, 10.0)):
        out = np.zeros(16, F32)
        assert lib.vcl_projection_matrix(size, aspect, n, f, P(out)) == 0
        r, t = size * aspect, size
        l, b = -r, -t
        ref = np.array([[2 * n / (r - l), 0, (r + l) / (r - l), 0], [0, 2 * n / (t - b), (t + b) / (t - b), 0],
                        [0, 0, -(f + n) / (f - n), -2 * f * n / (f - n)], [0, 0, -1, 0]])
        # the reference's sx = (znear / size) / aspect: the frustum's half-width is size * aspect
        np.testing.assert_allclose(m4(out), ref, rtol=2e-6, atol=1e-7)
        # near and far planes map to NDC z = -1 and +1
        for z, want in ((-n, -1.0), (-f, 1.0)):
            c = m4(out) @ np.array([0.0, 0.0, z, 1.0])
            assert abs(c[2] / c[3] - want) < 1e-5


def test_camera_matrices_and_backward(lib):
    rng = np.random.default_rng(5)
    q = unit_quats(1, 6)[0]
    eye = np.array([0.3, -0.2, 3.0], F32)
    R16 = flat(rot(lib, unit_quats(1, 7)[0]))
    V16, P16, M16 = (np.zeros(16, F32) for _ in range(3))
    assert lib.vcl_camera_matrices(P(q), P(eye), 0.5, 1.0, 1.0, 10.0, P(R16), P(V16), P(P16), P(M16)) == 0
    Rm, Vm, Pm = m4(R16), m4(V16), m4(P16)
    np.testing.assert_allclose(m4(M16), Pm @ Vm @ Rm, rtol=1e-5, atol=1e-6)
    G = rng.normal(size=(4, 4))
    G16 = flat(G)
    outs = [np.zeros(16, F32) for _ in range(3)]
    assert lib.vcl_camera_backward_host(P(G16), P(R16), P(V16), P(P16), 1, *[P(o) for o in outs]) == 0
    gR, gV, gP = (m4(o) for o in outs)

    def loss(Rx, Vx, Px):
        return np.sum(G * (Px @ Vx @ Rx))
    h = 1e-6
    for which, g in ((0, gR), (1, gV), (2, gP)):
        fd = np.zeros((4, 4))
        for i in range(4):
            for j in range(4):
                mats = [Rm.copy(), Vm.copy(), Pm.copy()]
                mats[which][i, j] += h
                lp = loss(*mats)
                mats[which][i, j] -= 2 * h
                fd[i, j] = (lp - loss(*mats)) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=1e-4, atol=1e-4 * np.abs(fd).max())
    # reference mode: the products as matrix.cu:236-238 writes them
    assert lib.vcl_camera_backward_host(P(G16), P(R16), P(V16), P(P16), 0, *[P(o) for o in outs]) == 0
    np.testing.assert_allclose(m4(outs[0]), G @ Vm.T @ Pm.T, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(m4(outs[1]), Rm.T @ G @ Pm.T, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(m4(outs[2]), Rm.T @ Vm.T @ G, rtol=1e-4, atol=1e-5)


def test_rotation_backward_fd(lib):
    rng = np.random.default_rng(8)
    for q in unit_quats(5, 9):
        G = rng.normal(size=(4, 4))
        G[3, :] = 0
        G[:, 3] = 0
        gq = np.zeros(4, F32)
        assert lib.vcl_rotation_backward_host(P(q), P(flat(G)), 1, P(gq)) == 0

        def loss(qq):  # R(q) as a function of unnormalised q (getRotation's polynomial)
            x, y, z, w = qq
            R = np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (z * x + w * y)],
                          [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                          [2 * (z * x - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])
            return np.sum(G[:3, :3] * R)
        # the polynomial itself is pinned by scipy for unit q
        np.testing.assert_allclose(loss(q.astype(np.float64)), np.sum(G[:3, :3] * SR.from_quat(q.astype(np.float64)).as_matrix()), rtol=1e-5, atol=1e-5)
        h = 1e-6
        fd = np.array([(loss(q + h * e) - loss(q - h * e)) / (2 * h) for e in np.eye(4)])
        np.testing.assert_allclose(gq, fd, rtol=1e-4, atol=1e-5)
        ref = np.zeros(4, F32)
        assert lib.vcl_rotation_backward_host(P(q), P(flat(G)), 0, P(ref)) == 0
        np.testing.assert_allclose(ref, 0.5 * fd, rtol=1e-4, atol=1e-5)


def test_rotation_adam_recurrence(lib):
    rng = np.random.default_rng(10)
    st = L.RotationAdam(0, 0.9, 0.999, 0.9, 0.999, 1e-2, 1e-8)
    q = unit_quats(1, 11)[0].copy()
    m = np.zeros(4)
    v = np.zeros(4)
    qq = q.astype(np.float64)
    rm, rv = 0.9, 0.999
    for it in range(5):
        g = rng.normal(size=4).astype(F32)
        assert lib.vcl_rotation_adam_step(C.byref(st), P(q), P(g)) == 0
        rm *= 0.9
        rv *= 0.999
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        qq = qq - (m / (1 - rm)) * 1e-2 / (np.sqrt(v / (1 - rv)) + 1e-8)
        qq /= np.linalg.norm(qq)
        np.testing.assert_allclose(q, qq, rtol=1e-5, atol=1e-6)
        assert st.it == it + 1
