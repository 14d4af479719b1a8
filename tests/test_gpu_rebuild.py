"""LBVH quality and the automatic rebuild policy (tt_build TT_REFIT_AUTO):

  SAH cost of a two-triangle tree ............... closed form 1 + (A0 + A1) / A_root (boxes
                                                  carry the builder's padding, 1e-5 (1 + max|x|))
  cost invariants ................................ rebuilds of the same vertices and a refit to
                                                  unmoved vertices give the built cost bit for bit
  policy ......................................... a random walk of the vertices raises the refit
                                                  cost; ratio 1.25 rebuilds (counted) and the cost
                                                  drops back; the image after an automatic rebuild
                                                  equals an explicit rebuild's bit for bit; ratio 0
                                                  never rebuilds
  drop-in ........................................ rebuild_gas = 3 through the autograd module
"""
import numpy as np
import pytest

from diff_tritracer import capi, scenes
from helpers import GpuScene, to_dev

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-rebuild]", *a, flush=True)


def _area(lo, hi):
    d = np.maximum(hi - lo, 0.0)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def test_sah_cost_two_triangles(cuda):
    import torch

    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.5],
                  [2.0, 2.0, 1.0], [3.0, 2.5, 1.0], [2.5, 3.0, 2.0]], np.float32)
    verts = to_dev(V, cuda)
    cumsum = to_dev(np.array([0, 3], np.int32), cuda)
    h = capi.TracerHandle()
    h.build(verts, cumsum, 2)
    out = torch.zeros(1, device=cuda)
    h.bvh_cost(out)
    torch.cuda.synchronize()
    def pad(lo, hi):  # the builder's padding per axis: 1e-5 (1 + max(|lo|, |hi|)) on both sides
        p = 1e-5 * (1.0 + np.maximum(np.abs(lo), np.abs(hi)))
        return lo - p, hi + p
    b0, b1 = pad(V[:3].min(0), V[:3].max(0)), pad(V[3:].min(0), V[3:].max(0))
    root = (np.minimum(b0[0], b1[0]), np.maximum(b0[1], b1[1]))
    want = 1.0 + (_area(*b0) + _area(*b1)) / _area(*root)
    got = float(out.item())
    _say(f"two triangles: SAH cost {got:.7f}, closed form {want:.7f}")
    assert abs(got - want) <= 1e-6 * want


def test_cost_invariants_and_policy(cuda):
    import torch

    sc, ro, rd = scenes.small_config(3, 128, 128)
    g = GpuScene(sc, cuda)
    s = torch.cuda.current_stream().cuda_stream
    c = torch.zeros(3, device=cuda)
    g.h.bvh_cost(c[0:1], s)
    g.h.build(g.verts, g.cumsum, g.P, capi.BUILD, s)
    g.h.bvh_cost(c[1:2], s)
    g.h.build(g.verts, g.cumsum, g.P, capi.REFIT, s)
    g.h.bvh_cost(c[2:3], s)
    torch.cuda.synchronize()
    _say(f"built cost {c.tolist()}")
    assert c[0] == c[1] == c[2] and float(c[0]) > 1.0
    base = g.verts.clone()
    rng = np.random.Generator(np.random.PCG64(5))
    for ratio, expect_rebuild in ((0.0, False), (1.25, True)):
        g.verts.copy_(base)
        g.h.build(g.verts, g.cumsum, g.P, capi.BUILD, s)
        g.h.set_rebuild_policy(ratio)
        n0 = g.h.build_stats()
        costs = []
        for step in range(8):  # a random walk of every vertex: the refit boxes grow
            g.verts.add_(to_dev((0.02 * rng.standard_normal(sc.verts.shape)).astype(np.float32), cuda))
            g.h.build(g.verts, g.cumsum, g.P, capi.REFIT_AUTO, s)
            costs.append(g.h.build_stats()[4])
        n1 = g.h.build_stats()
        _say(f"ratio {ratio}: refit costs {[round(x, 3) for x in costs]}; automatic rebuilds {n1[2] - n0[2]}; "
             f"cost after last build {n1[3]:.3f}")
        if expect_rebuild:
            assert n1[2] > n0[2]
            assert costs[-1] <= 1.25 * n1[3]
            # an automatic rebuild renders exactly as an explicit one
            a = g.forward(ro, rd, log_cap=24)
            g.h.build(g.verts, g.cumsum, g.P, capi.BUILD, s)
            b = g.forward(ro, rd, log_cap=24)
            assert np.array_equal(a["color"], b["color"])
            for x, y in zip(a["log"], b["log"]):
                assert np.array_equal(x, y)
        else:
            assert n1[2] == n0[2]
            assert costs[-1] > 1.25 * n1[3]


def test_dropin_rebuild_gas_auto(cuda):
    import torch
    from diff_tritracer import Tracer, TracerSettings

    sc, ro, rd = scenes.small_config(4, 64, 64)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    verts = t(sc.verts).reshape(-1)
    tracer = Tracer("forward.ptx", "backward.ptx").set_options(k=16, cutoff=1e-3).set_rebuild_policy(1.1)
    P = sc.P

    def render(mode):
        st = TracerSettings(bg=torch.zeros(3, device=cuda), scale_modifier=1.0, sh_degree=3, debug=False,
                            rebuild_gas=mode, extra_params=torch.zeros(0, device=cuda))
        with torch.no_grad():
            return tracer(t(ro), t(rd), verts, t(sc.sigma), t(sc.npts), t(sc.cumsum), P, t(sc.opac),
                          torch.zeros(P, device=cuda), torch.zeros(P, device=cuda), t(sc.shs), None, st)[0]

    render(0)
    rng = np.random.Generator(np.random.PCG64(6))
    for _ in range(6):
        verts.add_(t((0.03 * rng.standard_normal(verts.shape)).astype(np.float32)))
        img_auto = render(3)
    stats = tracer.build_stats()
    _say(f"drop-in build stats {stats}")
    assert stats[2] >= 1
    img_build = render(0)
    assert torch.equal(img_auto, img_build)
