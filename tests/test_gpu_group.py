"""Shared launches over several requests (vcl_group_*): one group step runs every request's
step in one launch per kernel, each workgroup finding its request in a device-side work table.

- A group of requests of uneven size (meshes of 1,280 to 81,920 triangles, 64^2 to 160^2
  images, 256^2 and 1024^2 textures, 16 to 128 lobes) ends its first step exactly where each
  request's own vcl_fit_step ends: the same image and loss bit for bit (the forward is
  deterministic), the same gradients up to the float-atomic summation order (first Adam moments
  within 1e-3).  After a few more steps (eager and graph-replayed) the losses agree to 2e-3:
  float atomics make every run's summation order differ (as in the reference), and Adam
  amplifies the last-digit differences -- for the 81,920-triangle 128^2 request two runs of the
  same plain step already differ by 2e-4 after 12 steps, and per-stream, graph-replayed and
  grouped runs spread over 1 % (measured), so the comparison stays at few steps.
- Requests whose steps launch different kernels cannot form a group (an error, not a silent
  fallback).
- The server's grouped mode serves a mix to the same results as its per-stream mode.
"""
import numpy as np
import pytest

from helpers import assert_close
from vcl_amd import ops
from vcl_amd import serve as SV

pytestmark = pytest.mark.gpu

SPECS = [{"level": 4, "W": 96, "tex": 256, "K": 16, "seed": 11},
         {"level": 5, "W": 160, "tex": 1024, "K": 64, "seed": 12},
         {"level": 3, "W": 64, "tex": 256, "K": 128, "seed": 13},
         {"level": 6, "W": 128, "tex": 256, "K": 16, "seed": 14}]


def _requests(specs):
    out = []
    for sp in specs:
        sc, tsc = SV.build(sp)
        out.append((sc, SV.render_target(tsc)))
    return out


def _state(f, sc):
    return {"loss": float(f.read("loss", (1,))[0]), "image": f.read("image", (sc.H, sc.W, sc.C)),
            "m_pos": f.read("m_pos", (sc.V * 3,)), "m_tex": f.read("m_tex", (sc.tw * sc.th * sc.C,)),
            "m_amp": f.read("m_amp", (sc.K * sc.C,))}


def test_group_matches_own_steps(gpu):
    reqs = _requests(SPECS)
    fits = [ops.Fit(sc, target=t) for sc, t in reqs]
    sigs = {ops.step_signature(f) for f in fits}
    assert len(sigs) == 1, "these request sizes share one step signature"
    print("step signature", sigs, flush=True)
    grp = ops.Group(fits)
    print("launches per group step", grp.launches(), flush=True)
    grp.step()
    got = [_state(f, sc) for f, (sc, _) in zip(fits, reqs)]
    own = []
    for sc, t in reqs:
        f = ops.Fit(sc, target=t)
        f.step()
        own.append(_state(f, sc))
        f.close()
    for i, (g, o) in enumerate(zip(got, own)):
        assert g["loss"] == o["loss"], f"request {i}: loss {g['loss']} vs {o['loss']}"
        np.testing.assert_array_equal(g["image"], o["image"], err_msg=f"request {i} image")
        for k in ("m_pos", "m_tex", "m_amp"):
            assert_close(g[k], o[k], 1e-3, f"request {i} {k}", atol_frac=1e-4)
    # two more steps, the last replayed from the group's graph
    for i in range(2):
        grp.step(graph=i >= 1)
    losses = [float(f.read("loss", (1,))[0]) for f in fits]
    grp.close()
    for f in fits:
        f.close()
    for i, (sc, t) in enumerate(reqs):
        f = ops.Fit(sc, target=t)
        for _ in range(3):
            f.step()
        ref = float(f.read("loss", (1,))[0])
        f.close()
        print(f"request {i}: 3 steps grouped {losses[i]:.7e} own {ref:.7e}", flush=True)
        np.testing.assert_allclose(losses[i], ref, rtol=2e-3)


def test_group_rejects_different_steps(gpu):
    (sc, t), = _requests(SPECS[:1])
    a = ops.Fit(sc, target=t)
    b = ops.Fit(sc, target=t, op_by_op=True)   # separate NormalAxis / ViewAxis / SGSpecular launches
    assert ops.step_signature(a) != ops.step_signature(b)
    with pytest.raises(Exception):
        ops.Group([a, b])
    with pytest.raises(Exception):
        ops.Group([a, a])                      # a request appears once
    a.close()
    b.close()


def test_server_grouped_matches_streams(gpu):
    reqs = _requests(SPECS)
    out = {}
    for grouped in (False, True):
        srv = SV.Server(max_active=8, auto_release=False, steps_per_round=2, grouped=grouped)
        tickets = [srv.submit(sc, t, 2 + i) for i, (sc, t) in enumerate(reqs)]
        srv.drain()
        st = {s["ticket"]: s for s in srv.finished()}
        assert len(st) == len(reqs) and all(s["status"] == 0 for s in st.values())
        out[grouped] = [(st[tk]["iterations"], float(srv.read(tk, "loss", (1,))[0])) for tk in tickets]
        srv.close()
    for (it_a, la), (it_b, lb) in zip(out[False], out[True]):
        assert it_a == it_b
        np.testing.assert_allclose(lb, la, rtol=2e-3)
