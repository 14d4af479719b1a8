"""The learned depth-interval proposal on the GPU (bf16 matrix cores) against the FP64
oracle (oracle/proposal.py, bf16 rounding emulated at the same points):

  forward outputs ................. max |gpu - oracle| <= 5e-3, mean <= 2e-5 (outputs O(1); the
                                    max is a rare input rounded to the other bf16 neighbour)
  interval mode ................... bit-exact fp32 formula of the raw outputs
  loss / gradient (B = 256) ....... loss within 1e-3 relative, gradient cosine >= 0.99999 and
                                    relative L2 error <= 1e-3
  one Adam step ................... the batch the GPU draws (splitmix64 sampler), Adam applied
                                    to the GPU's own gradient of that batch: <= 1e-6 abs
  training ........................ 300 steps on tracer-emitted intervals lower the loss
"""
import math

import numpy as np
import pytest

import proposal as PR
from diff_tritracer import capi, scenes
from helpers import GpuScene, to_dev

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-proposal]", *a, flush=True)


def _problem(seed, B):
    rng = np.random.Generator(np.random.PCG64(seed))
    params = PR.pack({name: rng.normal(0, 1.0 / math.sqrt(shp[-1]), shp) if name[0] == "W" else
                      rng.normal(0, 0.1, shp) for name, shp in PR.SHAPES}).astype(np.float32)
    o = rng.normal(0, 3.0, (B, 3)).astype(np.float32)
    d = rng.normal(0, 1.0, (B, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    iv = np.empty((B, 4), np.float32)
    iv[:, 0] = rng.uniform(1.0, 6.0, B)
    iv[:, 1] = iv[:, 0] + rng.uniform(0.0, 1.0, B)
    iv[:, 2] = np.where(rng.uniform(size=B) < 0.5, iv[:, 1] + 0.1, np.inf)
    iv[:, 3] = 0.0
    iv[rng.uniform(size=B) < 0.2, :3] = np.inf
    return params, o, d, iv


def test_forward_matches_oracle(cuda):
    import torch

    params, o, d, _ = _problem(0, 1000)
    net = capi.Proposal(seed=1)
    net.set_params(params)
    out = torch.empty(1000, 2, device=cuda)
    net.forward(to_dev(o, cuda), to_dev(d, cuda), out, capi.PROPOSAL_RAW)
    torch.cuda.synchronize()
    y = out.cpu().numpy()
    ref, _ = PR.forward(params.astype(np.float64), PR.encode(o, d, 8.0))
    err = np.abs(y - ref)
    _say(f"forward |gpu - oracle| max {err.max():.3g} mean {err.mean():.3g} (|y| max {np.abs(ref).max():.3g})")
    assert err.max() <= 5e-3 and err.mean() <= 2e-5
    iv = torch.empty(1000, 2, device=cuda)
    net.forward(to_dev(o, cuda), to_dev(d, cuda), iv, capi.PROPOSAL_INTERVAL, margin_start=0.05, margin_end=0.1)
    torch.cuda.synchronize()
    iv = iv.cpu().numpy()
    ts = np.float32(8.0)
    exp0 = np.maximum(np.float32(0), (y[:, 0] - np.float32(0.05)) * ts)
    exp1 = (y[:, 1] + np.float32(0.1)) * ts
    assert np.array_equal(iv[:, 0], exp0) and np.array_equal(iv[:, 1], exp1)


def test_forward_matches_oracle_far_origins(cuda):
    # origins ~100 units out: 2^7 pi o / 8 spans ~800 revolutions, past the range the
    # hardware sine takes directly; the kernels remove whole revolutions first
    import torch

    params, o, d, _ = _problem(4, 1000)
    o = (o + np.array([100.0, -80.0, 120.0], np.float32)).astype(np.float32)
    net = capi.Proposal(seed=1)
    net.set_params(params)
    out = torch.empty(1000, 2, device=cuda)
    net.forward(to_dev(o, cuda), to_dev(d, cuda), out, capi.PROPOSAL_RAW)
    torch.cuda.synchronize()
    y = out.cpu().numpy()
    ref, _ = PR.forward(params.astype(np.float64), PR.encode(o, d, 8.0))
    scale = max(1.0, float(np.abs(ref).max()))
    err = np.abs(y - ref)
    _say(f"far origins: |gpu - oracle| max {err.max():.3g} mean {err.mean():.3g} (|y| max {scale:.3g})")
    assert err.max() <= 5e-3 * scale and err.mean() <= 1e-4 * scale


def test_loss_gradient_matches_oracle(cuda):
    import torch

    params, o, d, iv = _problem(1, 256)
    net = capi.Proposal(seed=2)
    net.set_params(params)
    g = torch.empty(net.n_params, device=cuda)
    loss = torch.empty(1, device=cuda)
    net.loss_grad(to_dev(o, cuda), to_dev(d, cuda), to_dev(iv, cuda), g, loss)
    torch.cuda.synchronize()
    g = g.cpu().numpy().astype(np.float64)
    lref, gref = PR.loss_grad(params.astype(np.float64), o, d, iv)
    cos = g @ gref / np.linalg.norm(g) / np.linalg.norm(gref)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    _say(f"loss gpu {loss.item():.6g} oracle {lref:.6g}; gradient cosine {cos:.6f} rel L2 {rel:.3g}")
    assert abs(loss.item() - lref) <= 1e-3 * abs(lref)
    assert cos >= 0.99999 and rel <= 1e-3
    for name, (a, b) in {"W4": (PR.N_PARAMS - 130, PR.N_PARAMS - 2), "b4": (PR.N_PARAMS - 2, PR.N_PARAMS)}.items():
        np.testing.assert_allclose(g[a:b], gref[a:b], rtol=2e-2, atol=2e-4 * np.abs(gref[a:b]).max(), err_msg=name)


def test_adam_step_uses_the_sampled_batch(cuda):
    import torch

    params, o, d, iv = _problem(3, 500)
    seed = 7
    net = capi.Proposal(seed=seed)
    net.set_params(params)
    O, D, IV = to_dev(o, cuda), to_dev(d, cuda), to_dev(iv, cuda)
    hist = torch.zeros(1, device=cuda)
    net.train(O, D, IV, batch=64, steps=1, loss_hist=hist)
    torch.cuda.synchronize()
    after = net.get_params().astype(np.float64)
    idx = PR.sampled_rays(seed, 1, 64, 500)
    ref = capi.Proposal(seed=0)
    ref.set_params(params)
    g = torch.empty(ref.n_params, device=cuda)
    loss = torch.empty(1, device=cuda)
    ref.loss_grad(to_dev(o[idx], cuda), to_dev(d[idx], cuda), to_dev(iv[idx], cuda), g, loss)
    torch.cuda.synchronize()
    g = g.cpu().numpy().astype(np.float64)
    z = np.zeros(PR.N_PARAMS)
    exp, _, _ = PR.adam(params.astype(np.float64), z, z, g, 1)
    _say(f"one Adam step: max |gpu - numpy Adam(gpu grad)| {np.abs(after - exp).max():.3g}; loss {hist.item():.5g}")
    assert np.abs(after - exp).max() <= 1e-6
    assert hist.item() == pytest.approx(loss.item(), rel=1e-6)


def test_training_on_emitted_intervals_lowers_the_loss(cuda):
    import torch

    sc = scenes.uniform_scene(40_000, 6, offset_std=0.08)
    g = GpuScene(sc, cuda)
    os_, ds_, ivs = [], [], []
    for i in range(4):
        ro, rd = scenes.look_at_rays(64, 64, scenes.orbit_eye(3.0, 11, i), fov_deg=50.0)
        out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", interval_out=True)
        os_.append(ro.reshape(-1, 3))
        ds_.append(rd.reshape(-1, 3))
        ivs.append(out["interval"])
    O, D, IV = (to_dev(np.concatenate(a), cuda) for a in (os_, ds_, ivs))
    net = capi.Proposal(seed=4)
    hist = torch.zeros(300, device=cuda)
    net.train(O, D, IV, batch=1024, steps=300, loss_hist=hist)
    torch.cuda.synchronize()
    h = hist.cpu().numpy()
    _say(f"loss first 10 {h[:10].mean():.4f} last 30 {h[-30:].mean():.4f}")
    assert np.all(np.isfinite(h))
    assert h[-30:].mean() < 0.6 * h[:10].mean()


def test_trained_proposal_ablation(cuda):
    """Ablation on a scene where most rays turn opaque: the network, trained on 8 poses,
    bounds the held-out pose's forward. Cap only (exact continuation past the cap): the
    image equals the plain trace bit for bit. Proposed start + hard cap (fallback to the
    full trace only for rays with no hit in the interval, continuation for rays whose
    last blended hit is near the cap): lossy. The test records the node visits of both
    against the plain trace: the K-buffer already bounds each gather near the ray's
    last needed hit and the walk to the proposed start still descends the same upper
    levels, so neither mode saves traversal here."""
    import torch

    sc = scenes.uniform_scene(40_000, 6, offset_std=0.08)
    g = GpuScene(sc, cuda)
    os_, ds_, ivs = [], [], []
    for i in range(8):
        ro, rd = scenes.look_at_rays(96, 96, scenes.orbit_eye(3.0, 21, i), fov_deg=50.0)
        out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", interval_out=True)
        os_.append(ro.reshape(-1, 3))
        ds_.append(rd.reshape(-1, 3))
        ivs.append(out["interval"])
    O, D, IV = (to_dev(np.concatenate(a), cuda) for a in (os_, ds_, ivs))
    net = capi.Proposal(seed=5)
    net.train(O, D, IV, batch=2048, steps=1500)
    ro, rd = scenes.look_at_rays(96, 96, scenes.orbit_eye(3.0, 22, 0), fov_deg=50.0)
    R = 96 * 96
    ms, me = 0.05, 0.05
    cap = torch.empty(R, 2, device=cuda)
    net.forward(to_dev(ro, cuda), to_dev(rd, cuda), cap, capi.PROPOSAL_CAP, ms, me)
    both = torch.empty(R, 2, device=cuda)
    net.forward(to_dev(ro, cuda), to_dev(rd, cuda), both, capi.PROPOSAL_INTERVAL, ms, me)
    torch.cuda.synchronize()
    assert np.all(cap.cpu().numpy()[:, 0] == 0.0)

    def run(interval_in=None, hard=False):
        dev = cuda
        color, others = torch.zeros(3, R, device=dev), torch.zeros(7, R, device=dev)
        maxb, dens, scal = (torch.zeros(g.P, device=dev) for _ in range(3))
        trav = torch.zeros(R, 2, dtype=torch.int32, device=dev)
        a = capi.fwd_args(96, 96, to_dev(ro, dev), to_dev(rd, dev), g.verts, g.npts, g.cumsum, g.shs, 3, g.opac,
                          g.sigma, torch.zeros(3, device=dev), color, others, maxb, dens, scal, k=16, cutoff=1e-3,
                          chunk_mode="exact", interval_in=interval_in, trav_stats=trav, hard_cap=hard,
                          fallback_start=0.0, fallback_end=8.0 * me)
        g.h.fwd(a, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return color.cpu().numpy(), int(trav[:, 0].to(torch.int64).sum())

    c0, n0 = run()
    c1, n1 = run(cap)
    c2, n2 = run(both, hard=True)
    # the ceiling: each ray's own emitted interval, trusted (hard cap, rays that blend nothing
    # start past the far plane and are not traced)
    iv = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", interval_out=True)["interval"]
    first = np.where(np.isfinite(iv[:, 0]), iv[:, 0], iv[:, 2])
    stop = np.where(np.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    own = torch.from_numpy(np.stack([np.where(np.isfinite(first), first, 3.0e38),
                                     np.where(np.isfinite(stop), stop, 0.0)], 1).astype(np.float32)).to(cuda)
    c3, n3 = run(own, hard=True)
    mse = float(((c2 - c0) ** 2).mean())
    p = 10 * math.log10(1.0 / mse) if mse > 0 else float("inf")
    _say(f"held-out pose: BVH node visits/ray plain {n0 / R:.1f}, proposed cap {n1 / R:.1f} ({n1 / n0:.3f}x), "
         f"proposed start + hard cap {n2 / R:.1f} ({n2 / n0:.3f}x, PSNR vs plain {p:.2f} dB), own intervals "
         f"{n3 / R:.1f} ({n3 / n0:.3f}x)")
    # lossless modes render plain bit for bit
    assert np.array_equal(c0, c1) and np.array_equal(c0, c3)
    # what an interval can remove: the ray's own interval cuts the walk; the trained network's
    # start + hard cap removes traversal work and stays above 25 dB against plain. The proposed
    # cap alone can add visits: a ray still open at its cap walks again past it.
    assert n3 < 0.95 * n0  # measured 0.913x
    assert n2 <= n0
    assert p > 25.0


def test_forward_waits_for_intervals_made_on_another_stream(cuda):
    """interval_ready_event: the proposal network runs on a second stream while the forward
    forms its window records and probes its tiles; the forward's traversal waits for the
    event. Three rounds, each checked bit for bit against the same network and forward run
    one after the other on one stream (exact cap mode: the image is the plain trace's)."""
    import torch

    sc, ro, rd = scenes.small_config(0, 128, 128)
    g = GpuScene(sc, cuda)
    R = 128 * 128
    net = capi.Proposal(seed=7)
    o, d = to_dev(ro, cuda), to_dev(rd, cuda)
    main = torch.cuda.current_stream()
    s2 = torch.cuda.Stream()
    fork, ready = torch.cuda.Event(), torch.cuda.Event()
    ready.record(s2)

    def render(interval_in, event=None):
        color, others = torch.zeros(3, R, device=cuda), torch.zeros(7, R, device=cuda)
        maxb, dens, scal = (torch.zeros(g.P, device=cuda) for _ in range(3))
        a = capi.fwd_args(128, 128, o, d, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma,
                          torch.zeros(3, device=cuda), color, others, maxb, dens, scal, k=16, cutoff=1e-3,
                          chunk_mode="exact", interval_in=interval_in, interval_event=event)
        g.h.fwd(a, main.cuda_stream)
        return color

    for rnd in range(3):
        cap_serial = torch.empty(R, 2, device=cuda)
        net.forward(o, d, cap_serial, capi.PROPOSAL_CAP, 0.05, 0.05, stream=main.cuda_stream)
        c_serial = render(cap_serial)
        cap_over = torch.full((R, 2), float("nan"), device=cuda)
        fork.record(main)
        s2.wait_event(fork)
        net.forward(o, d, cap_over, capi.PROPOSAL_CAP, 0.05, 0.05, stream=s2.cuda_stream)
        ready.record(s2)
        c_over = render(cap_over, ready)
        torch.cuda.synchronize()
        same_cap = np.array_equal(cap_serial.cpu().numpy().view(np.uint32), cap_over.cpu().numpy().view(np.uint32))
        same_img = np.array_equal(c_serial.cpu().numpy().view(np.uint32), c_over.cpu().numpy().view(np.uint32))
        print(f"[gpu-proposal] round {rnd}: intervals equal {same_cap}, image equal {same_img}", flush=True)
        assert same_cap and same_img
        torch.randn(1000, device=cuda).sum().item()  # eager work between rounds
