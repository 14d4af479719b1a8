"""Fused ray generation and loss gradient in the persistent multi-camera launches.

  rays from the camera table (no ray buffers) ...... colour / others / statistics bit-identical
                                                     to the same launch fed tt_camera_rays-equal
                                                     host rays; colour <= 1e-6 vs the oracle's
                                                     FP32 forward camera by camera; gradients
                                                     within 1e-3 (FP32 atomics add in any order)
  fused L2 / L1 loss ................................ dL/dcolour and the loss against torch
                                                     autograd of the same mean loss on the
                                                     forward's colour (<= 1e-6 relative); the
                                                     loss bit-identical on every run
  one training step in a hipGraph (rays + forward +
  loss gradient + backward, no ray buffers) ......... three replays with eager work between
                                                     them, each equal to the eager fused step
  allocation-free forward with both fusions ........ bit-identical to tt_multi_forward
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import capi, scenes
from helpers import grad_close, oracle_scene, to_dev

pytestmark = pytest.mark.gpu

SIZES = [(96, 64), (128, 72), (80, 80), (150, 100)]


def _say(*a):
    print("[gpu-fused]", *a, flush=True)


@pytest.fixture(scope="module")
def setup(cuda):
    import torch

    sc = scenes.clustered_scene(150_000, 16, seed=7)
    w, h, ro, rd = scenes.multiview_rays(SIZES, seed=8)
    bases = scenes.multiview_bases(SIZES, seed=8)
    d = lambda a: to_dev(a, cuda)
    S = dict(verts=d(sc.verts), npts=d(sc.npts), cumsum=d(sc.cumsum), shs=d(sc.shs), opac=d(sc.opac.reshape(-1)),
             sigma=d(sc.sigma.reshape(-1)), ro=d(ro), rd=d(rd), bg=torch.tensor([0.1, 0.2, 0.3], device=cuda),
             cams=capi.camera_table(bases, cuda))
    tr = capi.TracerHandle()
    tr.build(S["verts"], S["cumsum"], sc.P)
    Rt = ro.shape[0]
    rng = np.random.Generator(np.random.PCG64(9))
    S["dpix"] = d(rng.standard_normal(3 * Rt).astype(np.float32))
    S["doth"] = torch.zeros(7 * Rt, device=cuda)
    S["target"] = d(rng.uniform(0.0, 1.0, 3 * Rt).astype(np.float32))
    return sc, w, h, ro, rd, S, tr, Rt


def _bufs(torch, cuda, P, Rt):
    return dict(color=torch.zeros(3 * Rt, device=cuda), others=torch.zeros(7 * Rt, device=cuda),
                maxb=torch.zeros(P, device=cuda), dens=torch.zeros(P, device=cuda), scal=torch.zeros(P, device=cuda),
                stats=torch.zeros(Rt, 2, dtype=torch.int32, device=cuda),
                dverts=torch.zeros(3 * P, 3, device=cuda), dsig=torch.zeros(P, device=cuda),
                dop=torch.zeros(P, device=cuda), dsh=torch.zeros(P, 16, 3, device=cuda),
                grad=torch.zeros(3 * Rt, device=cuda), loss=torch.zeros(1, device=cuda))


def _args(setup, B, fused_rays, loss_kind=None, dpix=None):
    sc, w, h, ro, rd, S, tr, Rt = setup
    fused_loss = loss_kind is not None
    return capi.multi_args(w, h, None if fused_rays else S["ro"], None if fused_rays else S["rd"], S["verts"],
                           S["npts"], S["cumsum"], S["shs"], 3, S["opac"], S["sigma"], S["bg"], B["color"],
                           B["others"], B["maxb"], B["dens"], B["scal"],
                           dL_dpix=B["grad"] if fused_loss else (S["dpix"] if dpix is None else dpix),
                           dL_dothers=S["doth"], dverts=B["dverts"], dsigma=B["dsig"], dopac=B["dop"], dsh=B["dsh"],
                           k=16, cutoff=1e-3, chunk_mode="exact", block=128, tiles_per_fetch=1, zero_grads=True,
                           stats=B["stats"], cameras=S["cams"] if fused_rays else None,
                           target=S["target"] if fused_loss else None, loss_grad=B["grad"] if fused_loss else None,
                           loss=B["loss"] if fused_loss else None, loss_kind=loss_kind or "l2")


def _host(B):
    return {k: v.cpu().numpy() for k, v in B.items()}


def test_fused_rays_equal_ray_buffers_and_oracle(setup, cuda):
    import torch

    sc, w, h, ro, rd, S, tr, Rt = setup
    s = torch.cuda.current_stream().cuda_stream
    A, B = _bufs(torch, cuda, sc.P, Rt), _bufs(torch, cuda, sc.P, Rt)
    a, b = _args(setup, A, False), _args(setup, B, True)
    _say("multi-camera step from ray buffers, then from the camera table")
    tr.multi_fwd(a, s)
    tr.multi_bwd(a, s)
    tr.multi_fwd(b, s)
    tr.multi_bwd(b, s)
    torch.cuda.synchronize()
    x, y = _host(A), _host(B)
    for k in ("color", "others", "stats", "maxb", "dens"):
        assert np.array_equal(x[k], y[k]), k
    for k in ("dverts", "dsig", "dop", "dsh"):
        ok, worst = grad_close(y[k], x[k], rel=1e-4)
        assert ok, (k, worst)
    Sor = oracle_scene(sc)
    opts = O.make_opts(k=16, cutoff=1e-3, chunk_mode="exact", bg=(0.1, 0.2, 0.3))
    off = 0
    for c, (wc, hc) in enumerate(SIZES):
        R = wc * hc
        ref = O.forward_f32(Sor, ro[off:off + R], rd[off:off + R], opts)
        d32 = np.abs(y["color"][3 * off:3 * (off + R)].reshape(3, R) - ref["color"]).max()
        _say(f"  camera {c} ({wc}x{hc}): colour |gpu-f32| {d32:.3g}")
        assert d32 <= 1e-6
        off += R


@pytest.mark.parametrize("kind", ["l2", "l1"])
def test_fused_loss_matches_torch(setup, cuda, kind):
    import torch

    sc, w, h, ro, rd, S, tr, Rt = setup
    s = torch.cuda.current_stream().cuda_stream
    B = _bufs(torch, cuda, sc.P, Rt)
    b = _args(setup, B, True, loss_kind=kind)
    tr.multi_fwd(b, s)
    torch.cuda.synchronize()
    col = B["color"].clone().requires_grad_(True)
    loss = ((col - S["target"]) ** 2).mean() if kind == "l2" else (col - S["target"]).abs().mean()
    loss.backward()
    g_ref = col.grad.cpu().numpy()
    g = B["grad"].cpu().numpy()
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    lv, lr = float(B["loss"].item()), float(loss.item())
    _say(f"{kind}: loss fused {lv:.8g} torch {lr:.8g}; dL/dcolour max rel err {err:.3g}")
    assert err <= 1e-6
    assert abs(lv - lr) <= 1e-5 * abs(lr)
    first = B["loss"].clone()
    tr.multi_fwd(b, s)
    torch.cuda.synchronize()
    assert torch.equal(first, B["loss"])
    # the backward fed by the fused gradient equals the one fed torch's gradient
    tr.multi_bwd(b, s)
    C = _bufs(torch, cuda, sc.P, Rt)
    c = _args(setup, C, True, dpix=col.grad.detach().contiguous())
    tr.multi_fwd(c, s)
    tr.multi_bwd(c, s)
    torch.cuda.synchronize()
    for k in ("dverts", "dsig", "dop", "dsh"):
        ok, worst = grad_close(B[k].cpu().numpy(), C[k].cpu().numpy(), rel=1e-4)
        assert ok, (k, worst)


def test_fused_training_step_in_a_graph(setup, cuda):
    import torch

    sc, w, h, ro, rd, S, tr, Rt = setup
    s = torch.cuda.current_stream().cuda_stream
    E = _bufs(torch, cuda, sc.P, Rt)
    e = _args(setup, E, True, loss_kind="l2")
    tr.multi_fwd(e, s)
    tr.multi_bwd(e, s)
    torch.cuda.synchronize()
    ref = _host(E)
    G = _bufs(torch, cuda, sc.P, Rt)
    g = _args(setup, G, True, loss_kind="l2")
    tr.multi_capture(g, 3, s)
    for rep in range(3):
        _say(f"graph replay {rep}")
        for v in G.values():
            v.fill_(7)
        G["maxb"].zero_()
        G["dens"].zero_()
        G["scal"].zero_()
        tr.multi_replay(s)
        torch.cuda.synchronize()
        x = _host(G)
        for k in ("color", "others", "stats", "maxb", "dens", "grad", "loss"):
            assert np.array_equal(x[k], ref[k]), (rep, k)
        for k in ("dverts", "dsig", "dop", "dsh"):
            ok, worst = grad_close(x[k], ref[k], rel=1e-4)
            assert ok, (rep, k, worst)


def test_noalloc_forward_with_fused_rays_and_loss(setup, cuda):
    import torch

    sc, w, h, ro, rd, S, tr, Rt = setup
    s = torch.cuda.current_stream().cuda_stream
    A, B = _bufs(torch, cuda, sc.P, Rt), _bufs(torch, cuda, sc.P, Rt)
    tr.multi_fwd(_args(setup, A, True, loss_kind="l1"), s)
    tr.multi_fwd_noalloc(_args(setup, B, True, loss_kind="l1"), s)
    torch.cuda.synchronize()
    x, y = _host(A), _host(B)
    for k in ("color", "others", "stats", "maxb", "dens", "grad", "loss"):
        assert np.array_equal(x[k], y[k]), k
