"""Kernel schedules (vcl_set_tuning) and the autotuner.

Every tuning must give the oracle's results: raster tile side and bin chunk change only how
the work is cut (the raster stays bit-exact), texel-splat rows change only the float
summation order of texel gradients.  The autotuner leaves the request's state as it found it.
"""
import numpy as np
import pytest

from helpers import F32, assert_close, forward_inputs, o_tex_bwd
from test_gpu_parity import N, T, _check_raster
from vcl_amd import scene as S

pytestmark = pytest.mark.gpu

TILINGS = [(8, 64), (8, 256), (16, 128), (32, 64), (32, 256)]


@pytest.fixture
def tuned(gpu):
    yield gpu
    gpu.set_tuning(16, 256, 16)


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    clip = rng.uniform(-1.3, 1.3, (3 * n, 4)).astype(F32)
    clip[:, 3] = rng.uniform(0.5, 2.0, 3 * n)
    clip[:, 2] = rng.uniform(-1.5, 1.5, 3 * n) * clip[:, 3]
    clip[900:960, 3] = -clip[900:960, 3]
    return clip, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@pytest.mark.parametrize("tile,chunk", TILINGS)
def test_raster_every_tiling(tuned, tile, chunk):
    tuned.set_tuning(tile, chunk, 16)
    clip, idx = _soup(2500, 21)
    _check_raster(tuned, clip, idx, 97, 75, what=f"soup tile {tile} chunk {chunk}")
    sc = S.config1_scene(1, W=128)
    d = forward_inputs(sc)
    _check_raster(tuned, d["clip"], sc.idx, sc.W, sc.H, what=f"config1 tile {tile} chunk {chunk}")


@pytest.mark.parametrize("rows", [16, 32, 64])
def test_texture_backward_every_splat(tuned, rows):
    import torch
    tuned.set_tuning(16, 256, rows)
    sc = S.config1_scene(1, W=160)
    d = forward_inputs(sc)
    rng = np.random.default_rng(5)
    gout = rng.normal(size=(1, sc.H, sc.W, sc.C)).astype(F32)
    gt_ref, guv_ref, _ = o_tex_bwd(d["rast"], d["uvp"], d["uvDA"], sc.tex, sc.tw, sc.th, sc.C, sc.L, gout, 0)
    levels = tuned.level_views(T(sc.tex), sc.tw, sc.th, sc.C, sc.L)
    gflat = torch.zeros(sc.tex.size, device="cuda")
    glv = tuned.level_views(gflat, sc.tw, sc.th, sc.C, sc.L)
    guv = torch.zeros(1, sc.H, sc.W, 2, device="cuda")
    guvda = torch.zeros(1, sc.H, sc.W, 4, device="cuda")
    tuned.texture_backward(T(d["rast"]), T(d["uvp"]), T(d["uvDA"]), levels, T(gout), glv, guv, guvda, da_mode=0)
    assert_close(N(gflat), gt_ref, 1e-3, f"texel grads, splat rows {rows}", atol_frac=1e-4)
    assert_close(N(guv), guv_ref, 1e-3, f"uv grads, splat rows {rows}")


def test_autotune_restores_state(gpu):
    sc = S.config1_scene(1, W=128)
    tgt_f = gpu.Fit(S.config1_scene(2, W=128))
    tgt_f.set_target_from_render()
    target = tgt_f.read("image", (1, sc.H, sc.W, sc.C))
    tgt_f.close()
    a = gpu.Fit(sc, target=target)
    b = gpu.Fit(sc, target=target)
    a.step()
    b.step()
    pos0, tex0 = a.read("pos", (sc.V, 3)), a.read("texture", (sc.tex.size,))
    best = a.autotune(steps=1)
    print("autotuned", best)
    assert best["raster_tile"] in (8, 16, 32) and best["bin_chunk"] in (64, 128, 256) and best["splat_rows"] in (16, 32, 64)
    assert a.tuning() == {k: best[k] for k in ("raster_tile", "bin_chunk", "splat_rows")}
    assert np.array_equal(a.read("pos", (sc.V, 3)), pos0) and np.array_equal(a.read("texture", (sc.tex.size,)), tex0)
    # the next step of the tuned request matches an untuned copy of the same request
    a.step(graph=True)
    b.step()
    np.testing.assert_allclose(float(a.read("loss", (1,))[0]), float(b.read("loss", (1,))[0]), rtol=1e-5)
    a.close()
    b.close()
