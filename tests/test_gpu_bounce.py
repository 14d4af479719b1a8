"""The bounce variant on the GPU (per-triangle reflectivity / IOR, one reflected and one
refracted ray per pixel from the primary's expected surface, equirect environment light)
against the oracle (40k splats, 144x112 camera, exact chunk mode):

  colour / others ........ <= 2e-6 abs vs the oracle's FP32 path, <= 1e-5 vs its FP64 replay
  statistics ............. consumed / blended hits over all three rays identical
  both with the backward re-traversing and with it reading the forward's hit lists of the
  three segments (capacity 32)
  gradients (FP64) ....... opacity, sigma, SH, reflectivity and environment texels elementwise
                           within 1e-3 max(|ref|, 1e-3 max|ref|); vertices elementwise within
                           1e-3 max(|ref|, 1e-3 max|ref|) + 1e-5 S, S = the sum over the blended
                           hits touching the entry of each hit's largest vertex-gradient component
                           (the oracle's abs mode), and relative L2 error <= 1e-4. A hit on a tiny
                           triangle near one edge gives two vertices gradients of ~1e3 and the third
                           a residual of ~1e-3 formed from those large quantities, and hits' terms
                           can cancel across rays: FP32 evaluation error scales with S, not with the
                           entry; 1e-5 S is ~170 FP32 ulps of S (3 of 360,000 entries need it here).
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import capi, scenes
from helpers import GpuScene, grad_close, oracle_scene, to_dev

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-bounce]", *a, flush=True)


@pytest.fixture(scope="module", params=[0, 32], ids=["retrace", "hitlist32"])
def case(cuda, request):
    return run_case(cuda, request.param)


def run_case(cuda, hitlist_cap=0):
    import torch

    sc = scenes.uniform_scene(40_000, 6, offset_std=0.08)
    ro, rd = scenes.look_at_rays(144, 112, scenes.orbit_eye(3.0, 6), fov_deg=50.0)
    rng = np.random.Generator(np.random.PCG64(3))
    refl = rng.beta(2.0, 5.0, sc.P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, sc.P).astype(np.float32)
    env = np.exp(rng.normal(0.0, 0.5, (64, 128, 3))).astype(np.float32)
    g = GpuScene(sc, cuda)
    R = ro.shape[0] * ro.shape[1]
    d = lambda a: to_dev(a, cuda)
    B = dict(reflectivity=d(refl), ior=d(ior), envmap=d(env), aux=torch.zeros(R, 20, device=cuda),
             dreflectivity=torch.zeros(sc.P, device=cuda), denvmap=torch.zeros(64, 128, 3, device=cuda))
    color, others = torch.zeros(3, R, device=cuda), torch.zeros(7, R, device=cuda)
    maxb, dens, scal = (torch.zeros(sc.P, device=cuda) for _ in range(3))
    stats = torch.zeros(R, 2, dtype=torch.int32, device=cuda)
    s = torch.cuda.current_stream().cuda_stream
    RO, RD = d(ro), d(rd)
    hl = capi.hitlist_buffers(R, hitlist_cap, cuda, segments=3) if hitlist_cap else None
    _say(f"bounce forward (hit list capacity {hitlist_cap})")
    a = capi.fwd_args(ro.shape[1], ro.shape[0], RO, RD, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma,
                      torch.zeros(3, device=cuda), color, others, maxb, dens, scal, k=16, cutoff=1e-3,
                      chunk_mode="exact", variant="bounce", stats=stats, bounce=B, hitlist=hl)
    g.h.fwd(a, s)
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    doth[2:5] = 0.3 * rng.standard_normal((3, R)).astype(np.float32)
    dverts = torch.zeros(g.verts.shape[0], 3, device=cuda)
    dsig, dop = torch.zeros(sc.P, device=cuda), torch.zeros(sc.P, device=cuda)
    dsh = torch.zeros(sc.P, 16, 3, device=cuda)
    _say("bounce backward")
    b = capi.bwd_args(ro.shape[1], ro.shape[0], RO, RD, g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma, color,
                      others, d(dpix), d(doth), dverts, dsig, dop, dsh, k=16, cutoff=1e-3, chunk_mode="exact",
                      zero_grads=True, bounce=B, hitlist=hl)
    g.h.bwd(b, s)
    torch.cuda.synchronize()
    out = dict(color=color.cpu().numpy(), others=others.cpu().numpy(), stats=stats.cpu().numpy(),
               dverts=dverts.cpu().numpy(), dsigma=dsig.cpu().numpy(), dopac=dop.cpu().numpy(), dsh=dsh.cpu().numpy(),
               drefl=B["dreflectivity"].cpu().numpy(), denv=B["denvmap"].cpu().numpy())
    S = oracle_scene(sc)
    opts = O.make_opts(k=16, cutoff=1e-3, chunk_mode="exact", variant="bounce", refl=refl, ior=ior, env=env)
    ref = O.forward_f32(S, ro, rd, opts)
    log = O.Log(S, ro, rd, opts)
    return out, ref, log, refl, env, dpix, doth


def test_bounce_forward_vs_oracle(case):
    out, ref, log, refl, env, _, _ = case
    d32 = np.abs(out["color"] - ref["color"]).max()
    o32 = np.abs(out["others"] - ref["others"]).max()
    c64, o64 = log.forward_f64_bounce(refl, env)
    d64 = np.abs(out["color"] - c64).max()
    _say(f"colour |gpu-f32| {d32:.3g}, others {o32:.3g}, |gpu-f64| {d64:.3g}; colour range "
         f"{out['color'].min():.3f}..{out['color'].max():.3f}")
    assert d32 <= 2e-6 and o32 <= 1e-6
    assert d64 <= 1e-5 and np.abs(out["others"] - o64).max() <= 1e-5
    assert np.array_equal(out["stats"], ref["stats"])


def test_bounce_gradients_vs_oracle(case):
    out, ref, log, refl, env, dpix, doth = case
    rb = log.backward_f64_bounce(refl, env, dpix, doth)
    for key, rkey in (("dopac", "dopac"), ("dsigma", "dsigma"), ("dsh", "dsh"), ("drefl", "drefl"),
                      ("denv", "denv")):
        ok, worst = grad_close(out[key], rb[rkey])
        _say(f"  {key}: worst ratio {worst:.3g}")
        assert ok, key
    g, r = out["dverts"].reshape(-1).astype(np.float64), rb["dverts"].reshape(-1)
    S = log.backward_f64_bounce(refl, env, dpix, doth, vertex_abs=True)["dverts"].reshape(-1)
    base = 1e-3 * np.maximum(np.abs(r), 1e-3 * np.abs(r).max())
    err = np.abs(g - r)
    conditioned = int((err > base).sum())
    worst = float((err / (base + 1e-5 * S)).max())
    rel = np.linalg.norm(g - r) / np.linalg.norm(r)
    _say(f"  dverts: worst ratio to the bar {worst:.3g}; {conditioned} of {r.size} entries need the 1e-5 S term; "
         f"relative L2 error {rel:.3g}")
    assert worst <= 1.0 and rel <= 1e-4


def test_two_pass_forward_equals_single_launch(cuda):
    """The default bounce forward (primary pass, then the secondary rays binned by region and
    direction) against the single launch the library uses when a debug hit log is asked
    for: colour, others, the backward record (aux), statistics, hit lists and hit counts
    bit-identical."""
    import torch

    sc = scenes.uniform_scene(20_000, 9, offset_std=0.08)
    ro, rd = scenes.look_at_rays(96, 80, scenes.orbit_eye(3.0, 4), fov_deg=50.0)
    rng = np.random.Generator(np.random.PCG64(5))
    refl = rng.beta(2.0, 5.0, sc.P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, sc.P).astype(np.float32)
    env = np.exp(rng.normal(0.0, 0.5, (32, 64, 3))).astype(np.float32)
    g = GpuScene(sc, cuda)
    R = ro.shape[0] * ro.shape[1]
    d = lambda a: to_dev(a, cuda)
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    for logged in (False, True):
        B = dict(reflectivity=d(refl), ior=d(ior), envmap=d(env), aux=torch.zeros(R, 20, device=cuda))
        color, others = torch.zeros(3, R, device=cuda), torch.zeros(7, R, device=cuda)
        maxb, dens, scal = (torch.zeros(sc.P, device=cuda) for _ in range(3))
        stats = torch.zeros(R, 2, dtype=torch.int32, device=cuda)
        hl = capi.hitlist_buffers(R, 32, cuda, segments=3)
        hl[0].fill_(-7)
        log = None
        if logged:
            log = (torch.full((R, 8), -1, dtype=torch.int32, device=cuda), torch.zeros(R, 8, device=cuda),
                   torch.full((R, 8), -1, dtype=torch.int32, device=cuda), torch.zeros(R, dtype=torch.int32, device=cuda))
        _say(f"bounce forward, {'single launch (logged)' if logged else 'two passes'}")
        a = capi.fwd_args(ro.shape[1], ro.shape[0], d(ro), d(rd), g.verts, g.npts, g.cumsum, g.shs, 3, g.opac,
                          g.sigma, torch.zeros(3, device=cuda), color, others, maxb, dens, scal, k=16, cutoff=1e-3,
                          chunk_mode="exact", variant="bounce", stats=stats, bounce=B, hitlist=hl, log=log)
        g.h.fwd(a, s)
        torch.cuda.synchronize()
        counts = hl[1].cpu().numpy()
        hits = hl[0].cpu().numpy()
        # only the stored entries are defined: mask each pixel's list beyond its count
        idx = np.arange(32)[None, :, None]
        valid = idx < np.maximum(counts, 0)[:, None, :]
        outs.append(dict(color=color.cpu().numpy(), others=others.cpu().numpy(), aux=B["aux"].cpu().numpy(),
                         stats=stats.cpu().numpy(), counts=counts, hits=np.where(valid[..., None], hits, 0),
                         maxb=maxb.cpu().numpy()))
    a, b = outs
    for k in a:
        assert np.array_equal(a[k].view(np.uint32) if a[k].dtype == np.float32 else a[k],
                              b[k].view(np.uint32) if b[k].dtype == np.float32 else b[k]), k
    assert (a["counts"][1:] > 0).any()


def test_backward_without_a_forward_of_the_same_size(cuda):
    """A bounce backward whose tracer last ran a forward of another image size (no per-tile
    record regions: fixed rows per lane, also for the globally sorted secondary records)
    gives the gradients of the regular one-shot pair up to FP32 summation order (the parity
    bar, 1e-3 of max(|ref|, 1e-3 max|ref|): SH sums of many cancelling terms differ by a few
    1e-4 between the two record layouts)."""
    import torch

    out, _, _, _, _, dpix, doth = run_case(cuda, 32)
    sc = scenes.uniform_scene(40_000, 6, offset_std=0.08)
    ro, rd = scenes.look_at_rays(144, 112, scenes.orbit_eye(3.0, 6), fov_deg=50.0)
    rng = np.random.Generator(np.random.PCG64(3))
    refl = rng.beta(2.0, 5.0, sc.P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, sc.P).astype(np.float32)
    env = np.exp(rng.normal(0.0, 0.5, (64, 128, 3))).astype(np.float32)
    g = GpuScene(sc, cuda)
    R = ro.shape[0] * ro.shape[1]
    d = lambda a: to_dev(a, cuda)
    s = torch.cuda.current_stream().cuda_stream
    B = dict(reflectivity=d(refl), ior=d(ior), envmap=d(env), aux=torch.zeros(R, 20, device=cuda),
             dreflectivity=torch.zeros(sc.P, device=cuda), denvmap=torch.zeros(64, 128, 3, device=cuda))
    color, others = torch.zeros(3, R, device=cuda), torch.zeros(7, R, device=cuda)
    maxb, dens, scal = (torch.zeros(sc.P, device=cuda) for _ in range(3))
    hl = capi.hitlist_buffers(R, 32, cuda, segments=3)
    a = capi.fwd_args(ro.shape[1], ro.shape[0], d(ro), d(rd), g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma,
                      torch.zeros(3, device=cuda), color, others, maxb, dens, scal, k=16, cutoff=1e-3,
                      chunk_mode="exact", variant="bounce", bounce=B, hitlist=hl)
    g.h.fwd(a, s)
    # a forward of another size on the same tracer drops the first one's record regions
    o2, d2 = scenes.look_at_rays(40, 24, scenes.orbit_eye(3.0, 6), fov_deg=50.0)
    g.forward(o2, d2, k=16, cutoff=1e-3, chunk_mode="exact")
    dverts = torch.zeros(g.verts.shape[0], 3, device=cuda)
    dsig, dop = torch.zeros(sc.P, device=cuda), torch.zeros(sc.P, device=cuda)
    dsh = torch.zeros(sc.P, 16, 3, device=cuda)
    _say("bounce backward with fixed record rows")
    b = capi.bwd_args(ro.shape[1], ro.shape[0], d(ro), d(rd), g.verts, g.npts, g.cumsum, g.shs, 3, g.opac, g.sigma,
                      color, others, d(dpix), d(doth), dverts, dsig, dop, dsh, k=16, cutoff=1e-3, chunk_mode="exact",
                      zero_grads=True, bounce=B, hitlist=hl)
    g.h.bwd(b, s)
    torch.cuda.synchronize()
    got = dict(dverts=dverts.cpu().numpy(), dsigma=dsig.cpu().numpy(), dopac=dop.cpu().numpy(), dsh=dsh.cpu().numpy(),
               drefl=B["dreflectivity"].cpu().numpy(), denv=B["denvmap"].cpu().numpy())
    for k in got:
        ok, worst = grad_close(got[k], out[k], rel=1e-3)
        assert ok, (k, worst)
