"""Persistent multi-camera path (one launch for all cameras, tiles in the probe's order
forward and the forward's measured order backward, optionally with the forward's hit
list), its hipGraph replay and the autotuner, against per-camera one-shot launches
through the C ABI.

Bars: colour/others bit-exact and statistics identical to per-camera launches
(same per-pixel code, exact chunk mode); gradients = sum of the per-camera
gradients within 1e-3 (FP32 atomics add in a different order)."""
import numpy as np
import pytest

from diff_tritracer import capi, scenes
from helpers import grad_close, to_dev

pytestmark = pytest.mark.gpu

SIZES = [(96, 64), (128, 72), (80, 80), (150, 100)]


def _say(*a):
    print("[gpu-multi]", *a, flush=True)


@pytest.fixture(scope="module")
def mv(cuda):
    import torch

    sc = scenes.clustered_scene(200_000, 16, seed=3)
    w, h, ro, rd = scenes.multiview_rays(SIZES, seed=4)
    d = lambda a: to_dev(a, cuda)
    S = dict(verts=d(sc.verts), npts=d(sc.npts), cumsum=d(sc.cumsum), shs=d(sc.shs), opac=d(sc.opac.reshape(-1)),
             sigma=d(sc.sigma.reshape(-1)), ro=d(ro), rd=d(rd), bg=torch.tensor([0.1, 0.2, 0.3], device=cuda))
    h_ = capi.TracerHandle()
    h_.build(S["verts"], S["cumsum"], sc.P)
    Rt = ro.shape[0]
    rng = np.random.Generator(np.random.PCG64(8))
    S["dpix"] = d(rng.standard_normal(3 * Rt).astype(np.float32))
    doth = np.zeros(7 * Rt, np.float32)
    S["doth"] = d(doth)
    return sc, w, h, S, h_, Rt


def _bufs(torch, cuda, P, Rt, M):
    return dict(color=torch.zeros(3 * Rt, device=cuda), others=torch.zeros(7 * Rt, device=cuda),
                maxb=torch.zeros(P, device=cuda), dens=torch.zeros(P, device=cuda), scal=torch.zeros(P, device=cuda),
                stats=torch.zeros(Rt, 2, dtype=torch.int32, device=cuda),
                dverts=torch.zeros(3 * P, 3, device=cuda), dsig=torch.zeros(P, device=cuda),
                dop=torch.zeros(P, device=cuda), dsh=torch.zeros(P, M, 3, device=cuda))


def _per_camera(mv, cuda, k=16):
    import torch

    sc, w, h, S, tr, Rt = mv
    B = _bufs(torch, cuda, sc.P, Rt, 16)
    off = 0
    s = torch.cuda.current_stream().cuda_stream
    for wc, hc in zip(w, h):
        R = wc * hc
        sl = slice(off, off + R)
        col = B["color"][3 * off:3 * (off + R)]
        oth = B["others"][7 * off:7 * (off + R)]
        a = capi.fwd_args(wc, hc, S["ro"][sl], S["rd"][sl], S["verts"], S["npts"], S["cumsum"], S["shs"], 3, S["opac"],
                          S["sigma"], S["bg"], col, oth, B["maxb"], B["dens"], B["scal"], k=k, cutoff=1e-3,
                          chunk_mode="exact", stats=B["stats"][sl])
        tr.fwd(a, s)
        b = capi.bwd_args(wc, hc, S["ro"][sl], S["rd"][sl], S["verts"], S["npts"], S["cumsum"], S["shs"], 3,
                          S["opac"], S["sigma"], col, oth, S["dpix"][3 * off:3 * (off + R)],
                          S["doth"][7 * off:7 * (off + R)], B["dverts"], B["dsig"], B["dop"], B["dsh"], k=k,
                          cutoff=1e-3, chunk_mode="exact", zero_grads=False)
        tr.bwd(b, s)
        off += R
    torch.cuda.synchronize()
    return {k_: v.cpu().numpy() for k_, v in B.items()}


def _multi_args(mv, B, k=16, block=128, tpf=1, hitlist_cap=0):
    sc, w, h, S, tr, Rt = mv
    hl = None
    if hitlist_cap:
        import torch
        hl = capi.hitlist_buffers(Rt, hitlist_cap, S["ro"].device)
        B["_hits"], B["_hit_counts"], B["_hit_colors"] = hl
    return capi.multi_args(w, h, S["ro"], S["rd"], S["verts"], S["npts"], S["cumsum"], S["shs"], 3, S["opac"],
                           S["sigma"], S["bg"], B["color"], B["others"], B["maxb"], B["dens"], B["scal"],
                           dL_dpix=S["dpix"], dL_dothers=S["doth"], dverts=B["dverts"], dsigma=B["dsig"],
                           dopac=B["dop"], dsh=B["dsh"], k=k, cutoff=1e-3, chunk_mode="exact", block=block,
                           tiles_per_fetch=tpf, zero_grads=True, stats=B["stats"], hitlist=hl)


@pytest.fixture(scope="module")
def ref(mv, cuda):
    _say("per-camera reference launches")
    return _per_camera(mv, cuda)


def _check(out, ref, grads=True):
    assert np.array_equal(out["color"], ref["color"])
    assert np.array_equal(out["others"], ref["others"])
    assert np.array_equal(out["stats"], ref["stats"])
    assert np.array_equal(out["maxb"].view(np.uint32), ref["maxb"].view(np.uint32))
    assert np.array_equal(out["dens"], ref["dens"])
    if grads:
        for key in ("dop", "dsig", "dsh", "dverts"):
            ok, worst = grad_close(out[key], ref[key])
            assert ok, (key, worst)


@pytest.mark.parametrize("block,tpf,k,cap", [(128, 1, 16, 0), (64, 2, 8, 0), (256, 4, 32, 0), (128, 1, 16, 32),
                                             (64, 2, 16, 3), (256, 1, 8, 32)])
def test_multi_equals_per_camera(mv, ref, cuda, block, tpf, k, cap):
    import torch

    sc, w, h, S, tr, Rt = mv
    _say(f"persistent multi-camera block={block} tiles/fetch={tpf} k={k} hit list capacity={cap}")
    B = _bufs(torch, cuda, sc.P, Rt, 16)
    a = _multi_args(mv, B, k=k, block=block, tpf=tpf, hitlist_cap=cap)
    s = torch.cuda.current_stream().cuda_stream
    tr.multi_fwd(a, s)
    tr.multi_bwd(a, s)
    torch.cuda.synchronize()
    _check({k_: v.cpu().numpy() for k_, v in B.items() if not k_.startswith("_")}, ref)


def test_graph_replay_equals_eager(mv, ref, cuda):
    import torch

    sc, w, h, S, tr, Rt = mv
    _say("hipGraph capture + 3 replays (hit list on)")
    B = _bufs(torch, cuda, sc.P, Rt, 16)
    a = _multi_args(mv, B, hitlist_cap=32)
    s = torch.cuda.current_stream().cuda_stream
    tr.multi_capture(a, 3, s)
    for rep in range(3):
        for v in B.values():
            v.fill_(7)  # eager work between replays; the graph must overwrite / re-zero everything it owns
        B["maxb"].zero_()
        B["dens"].zero_()
        tr.multi_replay(s)
        torch.cuda.synchronize()
        _check({k_: v.cpu().numpy() for k_, v in B.items() if not k_.startswith("_")}, ref)


def test_autotuner_picks_a_measured_configuration(mv, ref, cuda):
    import torch

    sc, w, h, S, tr, Rt = mv
    _say("autotuner sweep")
    B = _bufs(torch, cuda, sc.P, Rt, 16)
    a = _multi_args(mv, B)
    s = torch.cuda.current_stream().cuda_stream
    res = tr.multi_autotune(a, ks=(8, 16), blocks=(64, 128), tpf=(1, 2), reps=2, stream=s)
    assert len(res) == 8 and all(v > 0 for v in res.values())
    best = min(res, key=res.get)
    assert (a.k, a.block, a.tiles_per_fetch) == best
    B2 = _bufs(torch, cuda, sc.P, Rt, 16)
    a2 = _multi_args(mv, B2, k=a.k, block=a.block, tpf=a.tiles_per_fetch)
    tr.multi_fwd(a2, s)
    tr.multi_bwd(a2, s)
    torch.cuda.synchronize()
    _check({k_: v.cpu().numpy() for k_, v in B2.items()}, ref)


def test_graph_goes_stale_after_rebuild(mv, ref, cuda):
    """A graph records the tracer's nodes, stack depth and schedule: a refit keeps it valid,
    a rebuild or a forward of another tile layout (backward-only graph) makes the replay fail
    with TT_ERR_STATE before anything is launched."""
    import torch

    sc, w, h, S, tr, Rt = mv
    s = torch.cuda.current_stream().cuda_stream
    B = _bufs(torch, cuda, sc.P, Rt, 16)
    a = _multi_args(mv, B)
    _say("graph staleness: capture, refit, replay")
    tr.multi_capture(a, 3, s)
    tr.build(S["verts"], S["cumsum"], sc.P, capi.REFIT, s)
    B["maxb"].zero_()
    tr.multi_replay(s)
    torch.cuda.synchronize()
    _check({k_: v.cpu().numpy() for k_, v in B.items() if not k_.startswith("_")}, ref)
    _say("graph staleness: rebuild, replay must fail")
    tr.build(S["verts"], S["cumsum"], sc.P, capi.BUILD, s)
    with pytest.raises(RuntimeError, match="stale"):
        tr.multi_replay(s)
    _say("graph staleness: backward-only graph, one-shot forward in between")
    tr.multi_fwd(a, s)
    tr.multi_capture(a, 2, s)
    tr.multi_replay(s)
    torch.cuda.synchronize()
    W0, H0 = w[0], h[0]
    col0 = torch.zeros(3 * W0 * H0, device=cuda)
    oth0 = torch.zeros(7 * W0 * H0, device=cuda)
    fa = capi.fwd_args(W0, H0, S["ro"][:W0 * H0], S["rd"][:W0 * H0], S["verts"], S["npts"], S["cumsum"], S["shs"], 3,
                       S["opac"], S["sigma"], S["bg"], col0, oth0, B["maxb"], B["dens"], B["scal"], k=16,
                       cutoff=1e-3, chunk_mode="exact")
    tr.fwd(fa, s)
    with pytest.raises(RuntimeError, match="stale"):
        tr.multi_replay(s)
    torch.cuda.synchronize()
