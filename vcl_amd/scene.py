"""Seeded synthetic scenes for the fit workloads (meshes, cameras, textures, SG lobes).

Everything here is host-side numpy and deterministic for a given seed.  Camera maths
follows the reference's host code so that the same matrices feed the GPU path and the CPU
checker:

* ``look_at_quat``  -- Camera::setCam (matrix.cu:182-192)
* ``rotation_matrix`` -- getRotation (matrix.cu:43-53), column-major 4x4
* ``view_matrix``   -- getView (matrix.cu:175-180)
* ``projection``    -- Camera::forward (matrix.cu:209-219), ``size = znear * tan(fov/2)``
* ``pv_inverse``    -- ViewAxis::forward (material.cu:111): inverse(mat3(P) * mat3(V))

All matrices are returned as 16 float32 values in column-major order (glm layout).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


# ----------------------------------------------------------------------------- meshes
def icosphere(subdiv: int):
    """Unit icosphere: V = 10*4^s + 2, T = 20*4^s.  Midpoints are cached per edge."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def uv_sphere(segments: int, rings: int):
    """Unit UV sphere with single-vertex poles: V = 2 + segments*(rings-1),
    T = 2*segments*(rings-1) - ... (caps: segments each, body: 2*segments*(rings-2)).
    Also returns per-vertex (u, v) = (j/segments, i/rings)."""
    verts = [(0.0, 1.0, 0.0)]
    uvs = [(0.5, 1.0)]
    for i in range(1, rings):
        th = math.pi * i / rings
        y, r = math.cos(th), math.sin(th)
        for j in range(segments):
            ph = 2 * math.pi * j / segments
            verts.append((r * math.cos(ph), y, r * math.sin(ph)))
            uvs.append((j / segments, 1.0 - i / rings))
    verts.append((0.0, -1.0, 0.0))
    uvs.append((0.5, 0.0))
    faces = []
    S = segments
    for j in range(S):
        faces.append((0, 1 + (j + 1) % S, 1 + j))
    for i in range(rings - 2):
        b0 = 1 + i * S
        b1 = 1 + (i + 1) * S
        for j in range(S):
            j1 = (j + 1) % S
            faces.append((b0 + j, b0 + j1, b1 + j))
            faces.append((b0 + j1, b1 + j1, b1 + j))
    last = len(verts) - 1
    b = 1 + (rings - 2) * S
    for j in range(S):
        faces.append((last, b + j, b + (j + 1) % S))
    return np.array(verts, np.float64), np.array(faces, np.uint32), np.array(uvs, np.float64)


def value_noise3(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 2.0) -> np.ndarray:
    """Seeded 3-D value noise in [-1, 1]: lattice values from a hashed table, smoothstep
    trilinear interpolation, frequency x2 and amplitude x0.5 per octave, normalised."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)

    def lattice(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791)
        return table[perm[np.mod(h, 4096)]]

    total = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        q = p * freq + 100.0
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        s = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1 - s[:, 2]
                    acc += wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def vertex_normals(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (float64 in, float64 out, unit length)."""
    n = np.zeros_like(v)
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-20)


def spherical_uv(unit: np.ndarray) -> np.ndarray:
    u = 0.5 + np.arctan2(unit[:, 2], unit[:, 0]) / (2 * math.pi)
    v = 0.5 + np.arcsin(np.clip(unit[:, 1], -1, 1)) / math.pi
    return np.stack([u, v], 1)


# ----------------------------------------------------------------------------- cameras
def look_at_quat(eye, center, up):
    """Camera::setCam (matrix.cu:182-192) in float32: quaternion (w, x, y, z) of the view
    rotation whose rows are s, u, f."""
    eye, center, up = (np.asarray(a, F32) for a in (eye, center, up))

    def nrm(x):
        return (x / F32(np.sqrt(F32(np.dot(x, x))))).astype(F32)
    f = nrm(eye - center)
    s = nrm(np.cross(up, f).astype(F32))
    u = nrm(np.cross(f, s).astype(F32))
    half = F32(0.5)
    w = half * np.sqrt(F32(s[0] + u[1] + f[2] + F32(1)))
    x = (half if f[1] > u[2] else -half) * np.sqrt(abs(F32(s[0] - u[1] - f[2] + F32(1))))
    y = (half if s[2] > f[0] else -half) * np.sqrt(abs(F32(-s[0] + u[1] - f[2] + F32(1))))
    z = (half if u[0] > s[1] else -half) * np.sqrt(abs(F32(-s[0] - u[1] + f[2] + F32(1))))
    return np.array([w, x, y, z], F32)


def rotation_matrix(q) -> np.ndarray:
    """getRotation (matrix.cu:43-53) as a column-major (4,4) float32 array M[col][row]."""
    w, x, y, z = (F32(a) for a in q)
    qw2, qx2, qy2, qz2 = w * w, x * x, y * y, z * z
    qwx, qwy, qwz = w * x, w * y, w * z
    qxy, qyz, qzx = x * y, y * z, z * x
    two = F32(2)
    cols = [
        [qw2 + qx2 - qy2 - qz2, two * (qxy + qwz), two * (qzx - qwy), F32(0)],
        [two * (qxy - qwz), qw2 - qx2 + qy2 - qz2, two * (qyz + qwx), F32(0)],
        [two * (qzx + qwy), two * (qyz - qwx), qw2 - qx2 - qy2 + qz2, F32(0)],
        [F32(0), F32(0), F32(0), F32(1)],
    ]
    return np.array(cols, F32)


def _mul(a, b):
    """Column-major product A*B of (4,4) arrays indexed [col][row], float32."""
    return (b.astype(F32) @ a.astype(F32)).astype(F32)


def view_matrix(q, eye) -> np.ndarray:
    v = rotation_matrix(q)
    e = np.array([-eye[0], -eye[1], -eye[2], 1.0], F32)
    v[3] = (v.T.astype(F32) @ e).astype(F32)
    return v


def projection(size, aspect, znear, zfar) -> np.ndarray:
    sy = F32(znear) / F32(size)
    sx = sy / F32(aspect)
    fn = F32(zfar) - F32(znear)
    return np.array([[sx, 0, 0, 0], [0, sy, 0, 0], [0, 0, -(F32(zfar) + F32(znear)) / fn, -1],
                     [0, 0, F32(-2) * F32(zfar) * F32(znear) / fn, 0]], F32)


@dataclass
class Camera:
    mat: np.ndarray      # 16 float32, P*V*R column-major
    pvinv: np.ndarray    # 9 float32, inverse(mat3(P)*mat3(V)) column-major
    proj: np.ndarray
    view: np.ndarray


def make_camera(eye, center=(0, 0, 0), up=(0, 1, 0), fov_deg=45.0, aspect=1.0, znear=0.1, zfar=10.0,
                rot=None) -> Camera:
    size = F32(znear) * F32(math.tan(math.radians(fov_deg) / 2))
    q = look_at_quat(eye, center, up)
    V = view_matrix(q, np.asarray(eye, F32))
    P = projection(size, aspect, znear, zfar)
    R = np.eye(4, dtype=F32) if rot is None else rot
    M = _mul(_mul(P, V), R)
    p3 = P[:3, :3].astype(np.float64)
    v3 = V[:3, :3].astype(np.float64)
    pv = (v3 @ p3)          # column-major product mat3(P)*mat3(V) in [col][row] storage
    inv = np.linalg.inv(pv.T).T  # inverse in the same storage
    return Camera(M.reshape(16).astype(F32), inv.reshape(9).astype(F32), P, V)


def fibonacci_sphere(n: int, radius: float):
    """n points on a sphere, golden-angle spiral, avoiding the exact poles."""
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    th = math.pi * (1 + math.sqrt(5)) * i
    return np.stack([r * np.cos(th), z, r * np.sin(th)], 1) * radius


# ----------------------------------------------------------------------------- textures
def mip_levels(w: int, h: int, requested: int) -> int:
    """Texture::init level count (buffer.cu:380-385): min(LSB(w|h)+1, 16, requested)."""
    v = w | h
    lsb = (v & -v).bit_length() - 1
    return max(1, min(min(lsb + 1, 16), requested))


def level_shapes(w, h, L):
    return [(h >> l, w >> l) for l in range(L)]


def build_mips_np(level0: np.ndarray, L: int) -> np.ndarray:
    """Flat float32 buffer of all levels (2x2 box mean, float32 arithmetic)."""
    out = [level0.astype(F32)]
    for _ in range(1, L):
        a = out[-1]
        s = (a[0::2, 0::2] + a[0::2, 1::2]) + a[1::2, 0::2]
        s = s + a[1::2, 1::2]
        out.append((s * F32(0.25)).astype(F32))
    return np.concatenate([x.reshape(-1) for x in out]).astype(F32)


def sg_lobes(K: int, C: int, rng: np.random.Generator, sharp_lo=1.0, sharp_hi=50.0):
    z = rng.uniform(-1, 1, K)
    th = rng.uniform(-math.pi, math.pi, K)
    c = np.sqrt(1 - z * z)
    axis = np.stack([c * np.cos(th), c * np.sin(th), z], 1).astype(F32)
    sharp = rng.uniform(sharp_lo, sharp_hi, K).astype(F32)
    amp = rng.uniform(0, 1, (K, C)).astype(F32)
    return axis, sharp, amp


# ----------------------------------------------------------------------------- scenes
@dataclass
class Scene:
    """One fit request: parameters (pos, texture, lobes) plus frozen inputs."""
    W: int
    H: int
    pos: np.ndarray        # (V,3) f32, optimised
    normals: np.ndarray    # (V,3) f32
    uv: np.ndarray         # (V,2) f32
    idx: np.ndarray        # (T,3) u32
    tex: np.ndarray        # flat f32, all levels
    tw: int
    th: int
    C: int
    L: int
    axis: np.ndarray       # (K,3)
    sharp: np.ndarray      # (K,)
    amp: np.ndarray        # (K,C)
    cams: list = field(default_factory=list)
    rough: float = 0.5
    ior: float = 1.45

    @property
    def V(self):
        return self.pos.shape[0]

    @property
    def T(self):
        return self.idx.shape[0]

    @property
    def K(self):
        return self.sharp.shape[0]


def _texture(rng, S, C, L):
    lvl0 = rng.uniform(0, 1, (S, S, C)).astype(F32)
    return build_mips_np(lvl0, L)


def config1_scene(seed: int, W: int = 256) -> Scene:
    """Icosphere level 3 (642 V / 1280 T) with N(0, 0.01) vertex noise, camera at distance 3,
    fov 45, 512^2 x 4 texture with 10 levels, 32 lobes."""
    rng = np.random.default_rng(seed)
    unit, faces = icosphere(3)
    pos = (unit + rng.normal(0, 0.01, unit.shape)).astype(F32)
    normals = unit.astype(F32)
    uv = spherical_uv(unit).astype(F32)
    S, C = 512, 4
    L = mip_levels(S, S, 10)
    tex = _texture(rng, S, C, L)
    axis, sharp, amp = sg_lobes(32, C, rng)
    cam = make_camera((0.0, 0.0, 3.0))
    return Scene(W, W, pos, normals, uv, faces, tex, S, S, C, L, axis, sharp, amp, [cam])


_MESH_CACHE: dict = {}


def displaced_sphere(seed: int, segments=320, rings=157, amplitude=0.05):
    key = (seed, segments, rings, amplitude)
    if key not in _MESH_CACHE:
        unit, faces, uvs = uv_sphere(segments, rings)
        r = 1.0 + amplitude * value_noise3(unit, seed, 3)
        v = unit * r[:, None]
        _MESH_CACHE[key] = (v, faces, uvs, vertex_normals(v, faces))
    return _MESH_CACHE[key]


def config2_scene(seed: int, W: int = 1024, S: int = 2048, K: int = 128) -> Scene:
    """Displaced sphere, 49,922 V / 99,840 T, 1024^2, 2048^2 x 4 texture with 12 levels,
    128 lobes, camera at distance 3, fov 45."""
    rng = np.random.default_rng(seed)
    v, faces, uvs, n = displaced_sphere(1000 + seed)
    C = 4
    L = mip_levels(S, S, 12)
    tex = _texture(rng, S, C, L)
    axis, sharp, amp = sg_lobes(K, C, rng)
    cam = make_camera((0.0, 0.0, 3.0))
    return Scene(W, W, v.astype(F32), n.astype(F32), uvs.astype(F32), faces, tex, S, S, C, L,
                 axis, sharp, amp, [cam])


def multiview_cameras(n: int, radius: float = 3.0):
    return [make_camera(tuple(p)) for p in fibonacci_sphere(n, radius)]


def config3_scene(seed: int, W: int = 512, S: int = 2048, K: int = 64, views: int = 64) -> Scene:
    """Batched multi-view request: the displaced-sphere mesh (49,922 V / 99,840 T) seen from
    `views` cameras on a Fibonacci sphere of radius 3 (fov 45) at W x W each, one shared
    S^2 x 4 texture (12 levels), K lobes."""
    sc = config2_scene(seed, W=W, S=S, K=K)
    sc.cams = multiview_cameras(views, 3.0)
    return sc

