"""The drop-in module on the GPU: diff_tritracer.Tracer (the reference's nn.Module and
autograd Function, __init__.py:54-283) forward and backward through torch autograd, against
the oracle (small config: 10k splats, seed 0, 256x256 camera, k = 16, cutoff 1e-3, compat
chunk mode), with the forward's hit cache off (the backward re-traverses, as the reference)
and on (the backward reads the forward's stored hits).

Bars:
  colour ..................... <= 1e-5 abs vs the oracle's FP64 replay
  gradients (as autograd returns them: vertices flattened, sigma / opacity [P, 1], SH [P, M, 3])
                               elementwise within 1e-3 max(|ref|, 1e-3 max|ref|) of the FP64 oracle
  hit cache on / off ......... identical images
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes
from helpers import grad_close, oracle_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-dropin]", *a, flush=True)


@pytest.fixture(scope="module")
def case():
    sc, ro, rd = scenes.small_config(0)
    S = oracle_scene(sc)
    log = O.Log(S, ro, rd, O.make_opts(k=16, cutoff=1e-3))
    return sc, ro, rd, log


def _run(cuda, sc, ro, rd, dpix, doth, hit_cache):
    import torch
    from diff_tritracer import Tracer, TracerSettings

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    P = sc.P
    verts = t(sc.verts).reshape(-1).requires_grad_(True)
    sigma = t(sc.sigma).requires_grad_(True)
    opac = t(sc.opac).requires_grad_(True)
    shs = t(sc.shs).requires_grad_(True)
    tracer = Tracer("forward.ptx", "backward.ptx").set_options(k=16, cutoff=1e-3, chunk_mode="compat",
                                                              hit_cache=hit_cache)
    settings = TracerSettings(bg=torch.zeros(3, device=cuda), scale_modifier=1.0, sh_degree=3, debug=False,
                              rebuild_gas=0, extra_params=torch.zeros(0, device=cuda))
    color, _time, _scaling, _density, others, _maxb = tracer(
        t(ro), t(rd), verts, sigma, t(sc.npts), t(sc.cumsum), P, opac, torch.zeros(P, device=cuda),
        torch.zeros(P, device=cuda), shs, None, settings)
    H, W = ro.shape[:2]
    loss = (color * t(dpix).reshape(3, H, W)).sum() + (others * t(doth).reshape(7, H, W)).sum()
    loss.backward()
    torch.cuda.synchronize()
    return dict(color=color.detach().cpu().numpy().reshape(3, -1), dverts=verts.grad.cpu().numpy(),
                dsigma=sigma.grad.cpu().numpy(), dopac=opac.grad.cpu().numpy(), dsh=shs.grad.cpu().numpy())


@pytest.mark.parametrize("hit_cache", [0, 32, 64])
def test_dropin_autograd_vs_oracle(cuda, case, hit_cache):
    sc, ro, rd, log = case
    R = ro.shape[0] * ro.shape[1]
    rng = np.random.Generator(np.random.PCG64(13))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    doth[2:5] = 0.3 * rng.standard_normal((3, R)).astype(np.float32)
    _say(f"Tracer forward + autograd backward, hit_cache={hit_cache}")
    out = _run(cuda, sc, ro, rd, dpix, doth, hit_cache)
    c64, _ = log.forward_f64()
    err = float(np.abs(out["color"] - c64).max())
    _say(f"  colour |gpu-f64| {err:.3g}")
    assert err <= 1e-5
    rb = log.backward_f64(dpix, doth)
    assert out["dsigma"].shape == (sc.P, 1) and out["dopac"].shape == (sc.P, 1)
    assert out["dverts"].shape == (sc.verts.size,)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(out[key], rb[key])
        _say(f"  {key}: worst ratio vs FP64 {worst:.3g}")
        assert ok, key


def test_dropin_hit_cache_same_image(cuda, case):
    sc, ro, rd, _ = case
    R = ro.shape[0] * ro.shape[1]
    dpix = np.ones((3, R), np.float32)
    doth = np.zeros((7, R), np.float32)
    a = _run(cuda, sc, ro, rd, dpix, doth, 0)
    b = _run(cuda, sc, ro, rd, dpix, doth, 32)
    assert np.array_equal(a["color"].view(np.uint32), b["color"].view(np.uint32))
