"""Whole fit steps at the measured workload sizes against the oracle.

- Plain 1024^2 request: displaced sphere 49,922 V / 99,840 T, 2048^2 x 4 texture with 12
  levels, 128 lobes.  Triangle ids bit-exact, rasterizer (u, v, z), image and loss within
  1e-5, every parameter gradient within 1e-3; each op's output and each gradient is also
  stated against the FP64 oracle (printed: max, 99.9th percentile, fraction within the bar)
  and must be no less accurate against it than the reference-semantics float path.
- Batched 64 views x 512^2 of the same mesh with a shared 2048^2 x 4 texture and 64 lobes.
  This size takes the paths small tests never reach: the 1,024-pixel SG forward workgroups
  (>= 4,194,304 pixels), the SG lobe backward at its 4,096-workgroup cap with many chunks per
  workgroup, and the per-vertex project backward over more than 4 views.
Parameter gradients are read from the first Adam step's moments (m = (1 - beta1) g)."""
import numpy as np
import pytest

import oracle as O
from helpers import F32, OracleStep, assert_close, assert_fp64_no_worse, fp64_report
from vcl_amd import scene as S

pytestmark = pytest.mark.gpu

GRADS = (("pos", "pos"), ("tex", "tex"), ("axis", "axis"), ("sharp", "sharp"), ("amp", "amp"))


def _gpu_target(gpu, sc):
    t = gpu.Fit(sc)
    t.set_target_from_render()
    img = t.read("image", (len(sc.cams), sc.H, sc.W, sc.C))
    t.close()
    return img


def _sizes(sc):
    return {"pos": sc.V * 3, "tex": sc.tw * sc.th * sc.C, "axis": sc.K * 3, "sharp": sc.K, "amp": sc.K * sc.C}


def _check_step(gpu, sc, target, fp64):
    nv = len(sc.cams)
    P = nv * sc.H * sc.W
    O.set_threads(16)
    print(f"\n[fullsize] {nv} view(s) x {sc.W}^2, V={sc.V} T={sc.T}, tex {sc.tw}^2 x {sc.C} L={sc.L}, K={sc.K}", flush=True)
    fit = gpu.Fit(sc, target=target)
    fit.step()
    print("[fullsize] GPU step done; oracle step ...", flush=True)
    orc = OracleStep(sc, target, fp64=fp64)
    print("[fullsize] oracle done", flush=True)
    rast_g = fit.read("rast", (P, 4))
    rast_o = orc.buf("rast", (P, 4))
    assert np.array_equal(rast_g[:, 3], rast_o[:, 3]), f"triangle ids differ at {int(np.sum(rast_g[:, 3] != rast_o[:, 3]))} pixels"
    cov = rast_o[:, 3] > 0
    print(f"[fullsize] ids bit-exact over {P} pixels ({int(cov.sum())} covered)", flush=True)
    assert_close(rast_g[:, :3], rast_o[:, :3], 1e-5, "rast (u, v, z)")
    img_g = fit.read("image", orc.image.shape)
    assert_close(img_g, orc.image, 1e-5, "image")
    loss_g = float(fit.read("loss", (1,))[0])
    assert abs(loss_g - orc.loss) <= 1e-5 * abs(orc.loss), (loss_g, orc.loss)
    sizes = _sizes(sc)
    grads_g = {}
    for nm, _ in GRADS:
        m_g = fit.read("m_" + nm, (sizes[nm],))
        assert_close(m_g, orc.keep["m_" + nm], 1e-3, f"grad {nm}", atol_frac=1e-4)
        grads_g[nm] = m_g.astype(np.float64) / 0.1
    # updated parameters: a first Adam step moves each one by about lr*sign(g)
    for nm, key, base in (("pos", "pos", sc.pos), ("axis", "axis", sc.axis), ("sharpness", "sharp", sc.sharp),
                          ("amplitude", "amp", sc.amp), ("texture", "tex", sc.tex)):
        got = fit.read(nm, base.shape)
        ref = orc.keep[key].reshape(base.shape)
        agree = np.mean(np.isclose(got, ref, rtol=1e-5, atol=1e-6))
        assert agree > 0.999, f"{nm}: only {agree:.5f} of the updated values agree"
    if fp64:
        print("[fullsize] errors against the FP64 oracle (normwise relative):", flush=True)
        C = sc.C
        fw = (("rast (u, v, z)", "rast", (P, 4), lambda a: a[:, :3]), ("interpolate uv", "uvp", (P, 2), None),
              ("interpolate uvDA", "uvDA", (P, 4), None), ("normal axis", "nrm", (P, 3), None), ("view axis", "view", (P, 3), None),
              ("texture", "diff", (P, C), None), ("SGSpecular amplitude", "sAmp", (P,), None),
              ("SGSpecular sharpness", "sSharp", (P,), None), ("SG diffuse env", "denv", (P, C), None),
              ("SG specular env", "senv", (P, C), None), ("SG colour", "color", (P, C), None), ("antialias image", "img", (P, C), None))
        gpu_names = {"img": "image"}
        for what, nm, shape, sel in fw:
            g = fit.read(gpu_names.get(nm, nm), shape)
            o, d = orc.buf(nm, shape), orc.buf64(nm, shape)
            if sel is not None:
                g, o, d = sel(g), sel(o), sel(d)
            assert_fp64_no_worse(what, *fp64_report(what, g, o, d), 1e-5)
        assert_fp64_no_worse("loss", *fp64_report("loss", np.array([loss_g]), np.array([orc.loss]), np.array([orc.loss64])), 1e-5)
        gclip_g = fit.read("gclip", (nv * sc.V, 4))
        r = fp64_report("clip-space grad", gclip_g, orc.buf("gclip", (nv * sc.V, 4)), orc.buf64("gclip", (nv * sc.V, 4)), tol=1e-3)
        assert_fp64_no_worse("clip-space grad", *r, 1e-3)
        for nm, onm in (("pos", "gpos"), ("tex", "gtex"), ("axis", "gaxis"), ("sharp", "gsharp"), ("amp", "gamp")):
            n = sizes[nm]
            d = orc.buf64(onm, (n,) if onm != "gtex" else (sum((sc.tw >> l) * (sc.th >> l) * C for l in range(sc.L)),))[:n]
            o = orc.keep["m_" + nm].astype(np.float64) / 0.1
            assert_fp64_no_worse(f"grad {nm}", *fp64_report(f"grad {nm}", grads_g[nm], o, d, tol=1e-3), 1e-3)
    orc.close()
    fit.close()


def test_config2_plain_step_fullsize(gpu, orc):
    sc = S.config2_scene(1)
    target = _gpu_target(gpu, S.config2_scene(2))
    _check_step(gpu, sc, target, fp64=True)


def test_config3_batched_step_fullsize(gpu, orc):
    sc = S.config3_scene(1)
    tg = S.config3_scene(2)
    target = _gpu_target(gpu, tg)
    _check_step(gpu, sc, target, fp64=False)
