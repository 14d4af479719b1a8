"""GPU parity at the four scale configurations, against the CPU oracle on the same seeded inputs.

  C1  1M clustered splats (64 clusters, seed 1), one 1920x1080 camera, k = 16, cutoff 1e-3
      full frame, compat chunk mode (the drop-in default):
        hit logs (first 32 consumed hits: prim, t bits, decision) .. bit-exact vs oracle FP32
        colour / others ............................................ <= 1e-6 abs vs oracle FP32
        consumed / blended hits, max_blending, density ............. identical; scaling rtol 2e-6
      every 4th pixel in x and y (129,600 rays):
        colour / others ............................................ <= 1e-5 abs vs oracle FP64
        gradients, re-traversing backward (compat) and hit-list
        backward (exact) ........................................... elementwise 1e-3 vs oracle FP64
  C3  bounce variant: 1M splats (seed 3), Beta(2,5) reflectivity, U[1,1.6] IOR, 1024x512
      lognormal environment, 1920x1080, exact mode, hit lists of the three segments:
        colour full frame <= 2e-6 vs FP32, statistics identical; on every 4th pixel colour
        <= 1e-5 vs FP64 and every gradient (incl. vertices, reflectivity, texels) elementwise
        within 1e-3 of FP64
  C2  2M splats (seed 2), 8 cameras 640x480 .. 1920x1080 (9,004,032 rays) through the
      persistent multi-camera launches, compared with the oracle camera by camera:
        colour / others <= 1e-6 vs FP32 and statistics identical on every full frame;
        max_blending / density vs the oracle's per-camera maxima; gradients of every 8th
        pixel of each camera (upstream gradient zero elsewhere) vs the summed FP64 oracle
  C4  3M splats (128 clusters, seed 4), 1920x1080: plain exact trace vs the oracle on every
      4th pixel (FP32 colour, statistics, FP64 colour, gradients); the learned proposal
      (trained on 16 poses) in cap mode and the ceiling (each ray's own interval, hard cap)
      render the plain image bit for bit; node visits of each are reported

The backward's gradients are sums over pixels, so the sampled comparisons run the GPU backward
over the full frame with the upstream gradient zero outside the sample, and the oracle's FP64
backward over the sampled rays only. Rays are independent, so the oracle's forward over a
sample equals its forward over the full frame at those pixels.
"""
import math

import numpy as np
import pytest

import oracle as O
from diff_tritracer import capi, scenes
from helpers import GpuScene, compare_hit_logs_fast, grad_close, oracle_scene, to_dev

pytestmark = pytest.mark.gpu
K, CUT = 16, 1e-3


def _say(*a):
    print("[gpu-scale]", *a, flush=True)


def _sample(H, W, s):
    ys, xs = np.arange(0, H, s), np.arange(0, W, s)
    return (ys[:, None] * W + xs[None, :]).reshape(-1)


def _upstream(R, pix, seed, normals=0.3):
    rng = np.random.Generator(np.random.PCG64(seed))
    dpix = np.zeros((3, R), np.float32)
    doth = np.zeros((7, R), np.float32)
    dpix[:, pix] = rng.standard_normal((3, pix.size))
    doth[2:5, pix] = normals * rng.standard_normal((3, pix.size))
    return dpix, doth


def _check_grads(gb, rb, keys=("dopac", "dsigma", "dsh", "dverts")):
    for key in keys:
        ok, worst = grad_close(gb[key], rb[key])
        _say(f"  {key}: worst ratio {worst:.3g}")
        assert ok, key


# ------------------------------------------------------------------------------------ C1
@pytest.fixture(scope="module")
def c1(cuda):
    sc, ro, rd = scenes.scale_config(seed=1)
    _say(f"C1: {sc.P} splats, {ro.shape[1]}x{ro.shape[0]}")
    return sc, ro, rd, GpuScene(sc, cuda), oracle_scene(sc)


def test_c1_full_frame_forward(c1):
    sc, ro, rd, g, S = c1
    H, W = ro.shape[:2]
    R = H * W
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="compat", log_cap=32)
    opts = O.make_opts(k=K, cutoff=CUT, chunk_mode="compat")
    ref = O.forward_f32(S, ro, rd, opts)
    bad = compare_hit_logs_fast(out["log"], O.Log(S, ro, rd, opts).export(), R)
    d32 = np.abs(out["color"] - ref["color"]).max()
    do32 = np.abs(out["others"] - ref["others"]).max()
    _say(f"C1 full frame: hit-log mismatching rays {bad} of {R}; rgb |gpu-f32| {d32:.3g}, others {do32:.3g}; "
         f"blended hits/ray {out['stats'][:, 1].mean():.2f}")
    assert bad == 0
    assert d32 <= 1e-6 and do32 <= 1e-6
    assert np.array_equal(out["stats"], ref["stats"])
    assert np.array_equal(out["max_blending"].view(np.uint32), ref["max_blending"].view(np.uint32))
    assert np.array_equal(out["density"], ref["density"])
    np.testing.assert_allclose(out["scaling"], ref["scaling"], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("mode,cap", [("compat", 0), ("exact", 64)], ids=["compat-retrace", "exact-hitlist"])
def test_c1_sample_fp64_and_gradients(c1, mode, cap):
    sc, ro, rd, g, S = c1
    H, W = ro.shape[:2]
    R = H * W
    pix = _sample(H, W, 4)
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode=mode, hitlist_cap=cap)
    opts = O.make_opts(k=K, cutoff=CUT, chunk_mode=mode)
    ros, rds = ro.reshape(-1, 3)[pix], rd.reshape(-1, 3)[pix]
    ref = O.forward_f32(S, ros, rds, opts)
    log = O.Log(S, ros, rds, opts)
    c64, o64 = log.forward_f64()
    d32 = np.abs(out["color"][:, pix] - ref["color"]).max()
    d64 = np.abs(out["color"][:, pix] - c64).max()
    do64 = np.abs(out["others"][:, pix] - o64).max()
    _say(f"C1 {mode}, {pix.size} sampled rays: rgb |gpu-f32| {d32:.3g}, |gpu-f64| {d64:.3g}, others {do64:.3g}")
    assert d32 <= 1e-6 and d64 <= 1e-5 and do64 <= 1e-5
    assert np.array_equal(out["stats"][pix], ref["stats"])
    dpix, doth = _upstream(R, pix, 17)
    gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=K, cutoff=CUT, chunk_mode=mode,
                    hitlist=out.get("_hitlist"))
    rb = log.backward_f64(dpix[:, pix], doth[:, pix])
    _check_grads(gb, rb)


# ------------------------------------------------------------------------------------ C3
def _bounce_inputs(P):
    rng = np.random.Generator(np.random.PCG64(3))
    refl = rng.beta(2.0, 5.0, P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, P).astype(np.float32)
    env = np.exp(rng.normal(0.0, 0.5, (512, 1024, 3))).astype(np.float32)
    return refl, ior, env


def test_c3_bounce_at_scale(cuda):
    import torch

    sc = scenes.clustered_scene(1_000_000, 64, seed=3)
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    W, H = 1920, 1080
    ro, rd = scenes.look_at_rays(W, H, eye)
    R = W * H
    refl, ior, env = _bounce_inputs(sc.P)
    _say(f"C3: {sc.P} splats, bounce variant, {W}x{H}, envmap 1024x512")
    g = GpuScene(sc, cuda)
    d = lambda a: to_dev(a, cuda)
    B = dict(reflectivity=d(refl), ior=d(ior), envmap=d(env), aux=torch.zeros(R, 20, device=cuda),
             dreflectivity=torch.zeros(sc.P, device=cuda), denvmap=torch.zeros(512, 1024, 3, device=cuda))
    color, others = torch.zeros(3, R, device=cuda), torch.zeros(7, R, device=cuda)
    maxb, dens, scal = (torch.zeros(sc.P, device=cuda) for _ in range(3))
    stats = torch.zeros(R, 2, dtype=torch.int32, device=cuda)
    s = torch.cuda.current_stream().cuda_stream
    RO, RD = d(ro), d(rd)
    hl = capi.hitlist_buffers(R, 64, cuda, segments=3)
    a = capi.fwd_args(W, H, RO, RD, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma, torch.zeros(3, device=cuda),
                      color, others, maxb, dens, scal, k=K, cutoff=CUT, chunk_mode="exact", variant="bounce",
                      stats=stats, bounce=B, hitlist=hl)
    g.h.fwd(a, s)
    pix = _sample(H, W, 4)
    dpix, doth = _upstream(R, pix, 23)
    dverts = torch.zeros(g.verts.shape[0], 3, device=cuda)
    dsig, dop = torch.zeros(sc.P, device=cuda), torch.zeros(sc.P, device=cuda)
    dsh = torch.zeros(sc.P, 16, 3, device=cuda)
    b = capi.bwd_args(W, H, RO, RD, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma, color, others, d(dpix),
                      d(doth), dverts, dsig, dop, dsh, k=K, cutoff=CUT, chunk_mode="exact", zero_grads=True, bounce=B,
                      hitlist=hl)
    g.h.bwd(b, s)
    torch.cuda.synchronize()
    col, oth, st = color.cpu().numpy(), others.cpu().numpy(), stats.cpu().numpy()
    S = oracle_scene(sc)
    opts = O.make_opts(k=K, cutoff=CUT, chunk_mode="exact", variant="bounce", refl=refl, ior=ior, env=env)
    ref = O.forward_f32(S, ro, rd, opts)
    d32, o32 = np.abs(col - ref["color"]).max(), np.abs(oth - ref["others"]).max()
    _say(f"C3 full frame: colour |gpu-f32| {d32:.3g}, others {o32:.3g}; hits consumed/blended per pixel "
         f"{st[:, 0].mean():.2f} / {st[:, 1].mean():.2f}")
    assert d32 <= 2e-6 and o32 <= 1e-6
    assert np.array_equal(st, ref["stats"])
    log = O.Log(S, ro.reshape(-1, 3)[pix], rd.reshape(-1, 3)[pix], opts)
    c64, o64 = log.forward_f64_bounce(refl, env)
    d64 = np.abs(col[:, pix] - c64).max()
    _say(f"C3 {pix.size} sampled pixels: colour |gpu-f64| {d64:.3g}, others {np.abs(oth[:, pix] - o64).max():.3g}")
    assert d64 <= 1e-5 and np.abs(oth[:, pix] - o64).max() <= 1e-5
    rb = log.backward_f64_bounce(refl, env, dpix[:, pix], doth[:, pix])
    gb = dict(dopac=dop.cpu().numpy(), dsigma=dsig.cpu().numpy(), dsh=dsh.cpu().numpy(),
              drefl=B["dreflectivity"].cpu().numpy(), denv=B["denvmap"].cpu().numpy())
    gb["dverts"] = dverts.cpu().numpy()
    _check_grads(gb, rb, ("dopac", "dsigma", "dsh", "drefl", "denv", "dverts"))


# ------------------------------------------------------------------------------------ C2
def test_c2_multiview_per_camera_vs_oracle(cuda):
    import torch

    sc, ws, hs, ray_o, ray_d = scenes.multiview_config(seed=2, P=2_000_000)
    Rt = ray_o.shape[0]
    _say(f"C2: {sc.P} splats, {len(ws)} cameras, {Rt} rays")
    g = GpuScene(sc, cuda)
    d = lambda a: to_dev(a, cuda)
    RO, RD = d(ray_o), d(ray_d)
    color, others = torch.zeros(3 * Rt, device=cuda), torch.zeros(7 * Rt, device=cuda)
    maxb, dens, scal = (torch.zeros(sc.P, device=cuda) for _ in range(3))
    stats = torch.zeros(Rt, 2, dtype=torch.int32, device=cuda)
    offs = np.concatenate([[0], np.cumsum([w * h for w, h in zip(ws, hs)])])
    # upstream gradient on every 8th pixel of every camera
    dpix = np.zeros(3 * Rt, np.float32)
    doth = np.zeros(7 * Rt, np.float32)
    samples = []
    for c, (w, h) in enumerate(zip(ws, hs)):
        R = w * h
        pix = _sample(h, w, 8)
        dp, do = _upstream(R, pix, 100 + c)
        dpix[3 * offs[c]:3 * offs[c + 1]] = dp.reshape(-1)
        doth[7 * offs[c]:7 * offs[c + 1]] = do.reshape(-1)
        samples.append((pix, dp[:, pix], do[:, pix]))
    dverts = torch.zeros(g.verts.shape[0], 3, device=cuda)
    dsig, dop = torch.zeros(sc.P, device=cuda), torch.zeros(sc.P, device=cuda)
    dsh = torch.zeros(sc.P, 16, 3, device=cuda)
    hl = capi.hitlist_buffers(Rt, 32, cuda)
    a = capi.multi_args(ws, hs, RO, RD, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma,
                        torch.zeros(3, device=cuda), color, others, maxb, dens, scal, dL_dpix=d(dpix),
                        dL_dothers=d(doth), dverts=dverts, dsigma=dsig, dopac=dop, dsh=dsh, k=K, cutoff=CUT,
                        chunk_mode="exact", block=128, tiles_per_fetch=1, zero_grads=True, stats=stats, hitlist=hl)
    s = torch.cuda.current_stream().cuda_stream
    g.h.multi_fwd(a, s)
    g.h.multi_bwd(a, s)
    torch.cuda.synchronize()
    col, oth, st = color.cpu().numpy(), others.cpu().numpy(), stats.cpu().numpy()
    S = oracle_scene(sc)
    opts = O.make_opts(k=K, cutoff=CUT, chunk_mode="exact")
    mb = np.zeros(sc.P, np.float32)
    dn = np.zeros(sc.P, np.float32)
    rb = None
    for c, (w, h) in enumerate(zip(ws, hs)):
        R = w * h
        sl = slice(offs[c], offs[c + 1])
        ref = O.forward_f32(S, ray_o[sl], ray_d[sl], opts)
        gc = col[3 * offs[c]:3 * offs[c + 1]].reshape(3, R)
        go = oth[7 * offs[c]:7 * offs[c + 1]].reshape(7, R)
        d32, o32 = np.abs(gc - ref["color"]).max(), np.abs(go - ref["others"]).max()
        _say(f"C2 camera {c} ({w}x{h}): colour |gpu-f32| {d32:.3g}, others {o32:.3g}")
        assert d32 <= 1e-6 and o32 <= 1e-6
        assert np.array_equal(st[sl], ref["stats"])
        mb = np.maximum(mb, ref["max_blending"])
        dn = np.maximum(dn, ref["density"])
        pix, dp, do = samples[c]
        log = O.Log(S, ray_o[sl][pix], ray_d[sl][pix], opts)
        r = log.backward_f64(dp, do)
        rb = r if rb is None else {k: rb[k] + r[k] for k in rb}
    assert np.array_equal(maxb.cpu().numpy().view(np.uint32), mb.view(np.uint32))
    assert np.array_equal(dens.cpu().numpy(), dn)
    gb = dict(dverts=dverts.cpu().numpy(), dsigma=dsig.cpu().numpy(), dopac=dop.cpu().numpy(), dsh=dsh.cpu().numpy())
    _check_grads(gb, rb)


# ------------------------------------------------------------------------------------ C4
def test_c4_plain_and_proposal_at_scale(cuda):
    import torch

    sc = scenes.clustered_scene(3_000_000, 128, seed=4)
    W, H = 1920, 1080
    R = W * H
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    ro, rd = scenes.look_at_rays(W, H, eye)
    _say(f"C4: {sc.P} splats, {W}x{H}")
    g = GpuScene(sc, cuda)
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", hitlist_cap=64, interval_out=True)
    pix = _sample(H, W, 4)
    S = oracle_scene(sc)
    opts = O.make_opts(k=K, cutoff=CUT, chunk_mode="exact")
    ros, rds = ro.reshape(-1, 3)[pix], rd.reshape(-1, 3)[pix]
    ref = O.forward_f32(S, ros, rds, opts)
    log = O.Log(S, ros, rds, opts)
    c64, _ = log.forward_f64()
    d32, d64 = np.abs(out["color"][:, pix] - ref["color"]).max(), np.abs(out["color"][:, pix] - c64).max()
    _say(f"C4 {pix.size} sampled rays: rgb |gpu-f32| {d32:.3g}, |gpu-f64| {d64:.3g}")
    assert d32 <= 1e-6 and d64 <= 1e-5
    assert np.array_equal(out["stats"][pix], ref["stats"])
    dpix, doth = _upstream(R, pix, 29)
    gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=K, cutoff=CUT, chunk_mode="exact",
                    hitlist=out["_hitlist"])
    _check_grads(gb, log.backward_f64(dpix[:, pix], doth[:, pix]))

    # the learned proposal: trained on 16 poses (480x270) of the same scene, applied to this frame
    dev = cuda
    os_, ds_, ivs = [], [], []
    for i in range(16):
        e = scenes.orbit_eye(float(3.5 + (i % 5) * 0.25), 404, i)
        tro, trd = scenes.look_at_rays(480, 270, e)
        t = g.forward(tro, trd, k=K, cutoff=CUT, chunk_mode="exact", interval_out=True, stats=False)
        os_.append(tro.reshape(-1, 3))
        ds_.append(trd.reshape(-1, 3))
        ivs.append(t["interval"])
    net = capi.Proposal(seed=3)
    net.train(*(to_dev(np.concatenate(x), dev) for x in (os_, ds_, ivs)), batch=16384, steps=2000)
    cap = torch.empty(R, 2, device=dev)
    net.forward(to_dev(ro, dev), to_dev(rd, dev), cap, capi.PROPOSAL_CAP, 0.15, 0.05)
    both = torch.empty(R, 2, device=dev)
    net.forward(to_dev(ro, dev), to_dev(rd, dev), both, capi.PROPOSAL_INTERVAL, 0.15, 0.05)
    iv = out["interval"]
    first = np.where(np.isfinite(iv[:, 0]), iv[:, 0], iv[:, 2])
    stop = np.where(np.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    perfect = np.stack([np.where(np.isfinite(first), first, 3.0e38), np.where(np.isfinite(stop), stop, 0.0)],
                       1).astype(np.float32)
    torch.cuda.synchronize()

    def run(interval_in=None, hard=False, fb=(0.0, 0.0)):
        o = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=interval_in, hard_cap=hard,
                      fallback=fb, stats=False, trav=True)
        return o["color"], float(o["trav"][:, 0].astype(np.int64).sum()) / R

    c0, n0 = run()
    c1, n1 = run(cap.cpu().numpy())
    c2, n2 = run(both.cpu().numpy(), hard=True, fb=(0.0, 8.0 * 0.05))
    c3, n3 = run(perfect, hard=True)
    mse = float(((c2 - c0) ** 2).mean())
    p2 = 10 * math.log10(1.0 / mse) if mse > 0 else float("inf")
    _say(f"C4 node visits per ray: plain {n0:.1f}, proposal cap {n1:.1f}, proposal start + hard cap {n2:.1f} "
         f"(PSNR vs plain {p2:.2f} dB), own intervals {n3:.1f}")
    assert np.array_equal(c0.view(np.uint32), out["color"].view(np.uint32))
    assert np.array_equal(c1.view(np.uint32), c0.view(np.uint32))
    assert np.array_equal(c3.view(np.uint32), c0.view(np.uint32))
    assert n3 < 0.8 * n0
