"""Serving and batching of fit requests over the native scheduler (vcl_server_* of
include/vcl.h).

Server wraps one vcl_server: submit(scene, target, iterations, arrival_s) queues a request;
drain() runs scheduling rounds until every request finished; finished() returns per-request
timing.  The host arrays of a request are kept alive here until the server has admitted it.

uneven_scenes() and poisson_requests() generate the uneven-scene batch and the served stream:
displaced icospheres with triangle counts log-uniform in [1k, 500k], resolutions from
{256, 512, 1024, 2048}^2, textures from {256, 1024, 4096}^2 x 4 with full mips, {16, 64, 128}
lobes, and 20-200 iterations per served request, all from a seed.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import ops
from . import scene as S

GiB = 1 << 30


class Server:
    FIFO, SRPT = 0, 1

    def __init__(self, hbm_budget=200 * GiB, max_active=8, use_graph=True, steps_per_round=2, auto_release=True,
                 policy=0, max_stepping=0, autotune_steps=0, grouped=False, group_max_pixels=0):
        self.cfg = L.ServerConfig(int(hbm_budget), int(max_active), 1 if use_graph else 0, int(steps_per_round),
                                  1 if auto_release else 0, int(policy), int(max_stepping), int(autotune_steps),
                                  1 if grouped else 0, int(group_max_pixels))
        h = C.c_void_p()
        L.check(L.load().vcl_server_create(C.byref(self.cfg), C.byref(h)), "server_create")
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            L.load().vcl_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset_clock(self):
        L.check(L.load().vcl_server_reset_clock(self.h), "server_reset_clock")

    def clock(self):
        return float(L.load().vcl_server_clock(self.h))

    def submit(self, scene, target, iterations, arrival_s=0.0, lr=1e-2):
        d, keep = ops.fit_desc(scene, target, lr=lr)
        t = C.c_uint64()
        L.check(L.load().vcl_server_submit(self.h, C.byref(d), int(iterations), float(arrival_s), C.byref(t)), "server_submit")
        self._keep[t.value] = (d, keep)
        return t.value

    def progress(self):
        n = C.c_int()
        L.check(L.load().vcl_server_progress(self.h, C.byref(n)), "server_progress")
        return n.value

    def drain(self):
        L.check(L.load().vcl_server_drain(self.h), "server_drain")

    def finished(self):
        buf = (L.RequestStats * 256)()
        out = []
        while True:
            n = C.c_int()
            L.check(L.load().vcl_server_finished(self.h, buf, 256, C.byref(n)), "server_finished")
            for i in range(n.value):
                r = buf[i]
                out.append({"ticket": r.ticket, "arrival_s": r.arrival_s, "admit_s": r.admit_s, "finish_s": r.finish_s,
                            "iterations": r.iterations, "loss": r.final_loss, "device_bytes": r.device_bytes,
                            "status": r.status})
                self._keep.pop(r.ticket, None)
            if n.value < 256:
                return out

    def read(self, ticket, name, shape, dtype=np.float32):
        a = np.empty(shape, dtype)
        L.check(L.load().vcl_server_read(self.h, int(ticket), name.encode(), a.ctypes.data_as(C.c_void_p), a.nbytes), "server_read")
        return a

    def release(self, ticket):
        L.check(L.load().vcl_server_release(self.h, int(ticket)), "server_release")

    def tuning_table(self):
        """[(size-class key, (raster_tile, bin_chunk, splat_rows))] of the autotuner."""
        keys = (C.c_uint64 * 256)()
        ts = (L.Tuning * 256)()
        n = C.c_int()
        L.check(L.load().vcl_server_tunings(self.h, keys, ts, 256, C.byref(n)), "server_tunings")
        return [(keys[i], (ts[i].raster_tile, ts[i].bin_chunk, ts[i].splat_rows)) for i in range(n.value)]

    def load_tuning_table(self, table):
        n = len(table)
        keys = (C.c_uint64 * max(n, 1))(*[k for k, _ in table])
        ts = (L.Tuning * max(n, 1))(*[L.Tuning(*t) for _, t in table])
        L.check(L.load().vcl_server_load_tunings(self.h, keys, ts, n), "server_load_tunings")

    def tunings(self):
        """The tuning table keyed by a readable size class."""
        return {f"px~4^{k & 255} tri~8^{(k >> 8) & 255}": {"raster_tile": t[0], "bin_chunk": t[1], "splat_rows": t[2]}
                for k, t in self.tuning_table()}

    def memory(self):
        a, b, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        m = C.c_double()
        L.check(L.load().vcl_server_memory(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(m)), "server_memory")
        return {"in_use": a.value, "peak": b.value, "cached": c.value, "mean_in_use": m.value}


# ----------------------------------------------------------------------------- workloads
_ICO_T = {3: 1280, 4: 5120, 5: 20480, 6: 81920, 7: 327680}
_MESHES: dict = {}
_TEXTURES: dict = {}


def displaced_icosphere(level: int, seed: int, amplitude=0.05):
    key = (level, seed)
    if key not in _MESHES:
        unit, faces = S.icosphere(level)
        r = 1.0 + amplitude * S.value_noise3(unit, seed, 3)
        v = unit * r[:, None]
        _MESHES[key] = (v.astype(np.float32), faces, S.spherical_uv(unit).astype(np.float32),
                        S.vertex_normals(v, faces).astype(np.float32))
    return _MESHES[key]


def _texture(size, seed):
    key = (size, seed)
    if key not in _TEXTURES:
        rng = np.random.default_rng(10_000 + seed)
        L_ = S.mip_levels(size, size, 16)
        _TEXTURES[key] = (S.build_mips_np(rng.uniform(0, 1, (size, size, 4)).astype(np.float32), L_), L_)
    return _TEXTURES[key]


def scene_of(level, W, tex, K, seed, tex_seed=0):
    """One request of the uneven mix: displaced icosphere of `level`, one camera at distance 3."""
    v, faces, uv, n = displaced_icosphere(level, 1000 + seed % 4)
    rng = np.random.default_rng(seed)
    texture, L_ = _texture(tex, tex_seed)
    axis, sharp, amp = S.sg_lobes(K, 4, rng)
    cam = S.make_camera((0.0, 0.0, 3.0))
    return S.Scene(W, W, v.copy(), n, uv, faces, texture, tex, tex, 4, L_, axis, sharp, amp, [cam])


def sample_size(rng):
    """Triangle count log-uniform in [1k, 500k] (rounded to the nearest icosphere level),
    resolution, texture size and lobe count of one request."""
    t = math.exp(rng.uniform(math.log(1000), math.log(500_000)))
    level = min(_ICO_T, key=lambda k: abs(math.log(_ICO_T[k]) - math.log(t)))
    W = int(rng.choice([256, 512, 1024, 2048]))
    tex = int(rng.choice([256, 1024, 4096]))
    K = int(rng.choice([16, 64, 128]))
    return level, W, tex, K


def uneven_scenes(n, seed=0, max_res=2048, max_tex=4096):
    """n independent requests of uneven size; each with a target rendered from a second
    parameter seed of the same mesh and camera (rendered on the GPU by the caller)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        level, W, tex, K = sample_size(rng)
        W, tex = min(W, max_res), min(tex, max_tex)
        out.append({"level": level, "W": W, "tex": tex, "K": K, "seed": seed * 1000 + i})
    return out


def poisson_requests(n, rate, seed=0, max_res=2048, max_tex=4096):
    """n served requests: uneven sizes, 20-200 iterations, Poisson arrivals at `rate` req/s."""
    rng = np.random.default_rng(seed + 7)
    reqs = uneven_scenes(n, seed, max_res, max_tex)
    t = 0.0
    for r in reqs:
        t += rng.exponential(1.0 / rate)
        r["arrival_s"] = t
        r["iterations"] = int(rng.integers(20, 201))
    return reqs


def build(spec):
    """(scene, target scene) of a request spec: the target uses a second texture/lobe seed."""
    sc = scene_of(spec["level"], spec["W"], spec["tex"], spec["K"], spec["seed"], tex_seed=0)
    tsc = scene_of(spec["level"], spec["W"], spec["tex"], spec["K"], spec["seed"] + 500_000, tex_seed=1)
    return sc, tsc


def render_target(tsc):
    f = ops.Fit(tsc)
    f.set_target_from_render()
    img = f.read("image", (1, tsc.H, tsc.W, tsc.C))
    f.close()
    return img


def percentile(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else float("nan")
