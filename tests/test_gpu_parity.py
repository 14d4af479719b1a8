"""GPU parity of the HIP tracer against the CPU oracle (small config, C ABI).

Small config: 10k triangle splats, seed 0, 256x256 camera at radius 3, k = 16,
transmittance cutoff 1e-3, upstream dL/dRGB ~ N(0,1).

Bars (fixed by the project, see the test bodies):
  hit lists (prim, t bits, decision) ........ bit-exact vs the oracle's FP32 path
  RGB / others ............................... <= 1e-6 abs vs oracle FP32, <= 1e-5 abs vs oracle FP64
  max_blending, density ...................... bit-exact vs oracle FP32
  gradients .................................. elementwise |g-ref| <= 1e-3 max(|ref|, 1e-3 max|ref|) vs FP64
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes
from helpers import GpuScene, compare_hit_logs, grad_close, oracle_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-parity]", *a, flush=True)


@pytest.fixture(scope="module")
def small(cuda):
    sc, ro, rd = scenes.small_config(0)
    return sc, ro, rd, GpuScene(sc, cuda), oracle_scene(sc)


@pytest.fixture(scope="module")
def small_fwd(small):
    sc, ro, rd, g, S = small
    _say("small forward, k=16, cutoff=1e-3")
    opts = O.make_opts(k=16, cutoff=1e-3)
    out = g.forward(ro, rd, k=16, cutoff=1e-3, log_cap=48)
    ref = O.forward_f32(S, ro, rd, opts)
    log = O.Log(S, ro, rd, opts)
    return out, ref, log


def test_hit_lists_bit_exact(small, small_fwd):
    out, ref, log = small_fwd
    R = out["color"].shape[1]
    bad = compare_hit_logs(out["log"], log.export(), R)
    _say("hit-log mismatching rays:", bad, "of", R)
    assert bad == 0
    assert np.array_equal(out["stats"], ref["stats"])


def test_forward_rgb_vs_oracle(small, small_fwd):
    out, ref, log = small_fwd
    d32 = np.abs(out["color"] - ref["color"]).max()
    do32 = np.abs(out["others"] - ref["others"]).max()
    c64, o64 = log.forward_f64()
    d64 = np.abs(out["color"].astype(np.float64) - c64).max()
    do64 = np.abs(out["others"].astype(np.float64) - o64).max()
    _say(f"rgb |gpu-f32| {d32:.3g} others {do32:.3g}; |gpu-f64| rgb {d64:.3g} others {do64:.3g}")
    assert d32 <= 1e-6 and do32 <= 1e-6
    assert d64 <= 1e-5 and do64 <= 1e-5
    assert np.all(out["others"][[0, 5, 6]] == 0.0)


def test_forward_statistics(small, small_fwd):
    out, ref, _ = small_fwd
    assert np.array_equal(out["max_blending"].view(np.uint32), ref["max_blending"].view(np.uint32))
    assert np.array_equal(out["density"], ref["density"])
    np.testing.assert_allclose(out["scaling"], ref["scaling"], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("with_normals", [False, True])
def test_backward_vs_oracle(small, small_fwd, with_normals):
    sc, ro, rd, g, S = small
    out, ref, log = small_fwd
    R = ro.shape[0] * ro.shape[1]
    rng = np.random.Generator(np.random.PCG64(7))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    if with_normals:
        doth[2:5] = rng.standard_normal((3, R)).astype(np.float32)
    _say("small backward, normals" if with_normals else "small backward")
    gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=16, cutoff=1e-3)
    rb = log.backward_f64(dpix, doth)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(gb[key], rb[key])
        _say(f"  {key}: worst ratio {worst:.3g}")
        assert ok, key


@pytest.mark.parametrize("k", [8, 32])
def test_exact_mode_k_invariant(small, k):
    sc, ro, rd, g, S = small
    _say("exact chunk mode, k =", k)
    a = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact")
    b = g.forward(ro, rd, k=k, cutoff=1e-3, chunk_mode="exact")
    assert np.array_equal(a["color"], b["color"])
    assert np.array_equal(a["others"], b["others"])
    ref = O.forward_f32(S, ro, rd, O.make_opts(k=k, cutoff=1e-3, chunk_mode="exact"))
    assert np.abs(b["color"] - ref["color"]).max() <= 1e-6


@pytest.mark.parametrize("variant,extra", [
    ("reflect", [0.6, 0.1, -0.2, 0.3]),
    ("refract", [0.6, 0.1, -0.2, 0.3]),
    ("envlight", [0.6, 0.1, -0.2, 0.3, 4.0]),
])
def test_variants_vs_oracle(small, variant, extra):
    sc, ro, rd, g, S = small
    _say("variant", variant)
    bg = (0.2, 0.3, 0.4)
    out = g.forward(ro, rd, k=16, cutoff=1e-4, variant=variant, extra=extra, bg=bg)
    opts = O.make_opts(k=16, cutoff=1e-4, variant=variant, extra=extra, bg=bg)
    ref = O.forward_f32(S, ro, rd, opts)
    d = np.abs(out["color"] - ref["color"])
    _say(f"  max |gpu-f32| {d.max():.3g}")
    assert d.max() <= 1e-6
    c64, _ = O.Log(S, ro, rd, opts).forward_f64()
    assert np.abs(out["color"] - c64).max() <= 1e-5
    assert np.array_equal(out["stats"], ref["stats"])


def test_refit_matches_rebuild(cuda):
    from diff_tritracer import capi

    sc, ro, rd = scenes.small_config(0, 128, 128)
    g = GpuScene(sc, cuda)
    rng = np.random.Generator(np.random.PCG64(3))
    moved = sc.verts + 0.01 * rng.standard_normal(sc.verts.shape).astype(np.float32)
    g.verts.copy_(g.torch.from_numpy(moved).to(cuda))
    _say("refit vs rebuild")
    g.h.build(g.verts, g.cumsum, g.P, capi.REFIT)
    a = g.forward(ro, rd, log_cap=32)
    g.h.build(g.verts, g.cumsum, g.P, capi.BUILD)
    b = g.forward(ro, rd, log_cap=32)
    assert np.array_equal(a["color"], b["color"])
    for x, y in zip(a["log"], b["log"]):
        assert np.array_equal(x, y)


def test_dense_scene_more_than_16_blended_hits(cuda):
    """More blended hits per pixel than the backward's default record pool holds (16 per
    lane): the first backward sends the overflow through global atomics, later ones get a
    pool grown to the forward's total; gradients match the FP64 oracle either way."""
    sc = scenes.uniform_scene(30_000, 8, offset_std=0.15)
    sc.opac *= 0.15
    ro, rd = scenes.look_at_rays(128, 96, scenes.orbit_eye(3.0, 8), fov_deg=40.0)
    g = GpuScene(sc, cuda)
    S = oracle_scene(sc)
    R = 128 * 96
    out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", hitlist_cap=128)
    mean_blended = out["stats"][:, 1].mean()
    _say(f"dense scene: {mean_blended:.1f} blended hits per pixel (max {out['stats'][:, 1].max()})")
    assert mean_blended > 16
    rng = np.random.Generator(np.random.PCG64(4))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    rb = O.Log(S, ro, rd, O.make_opts(k=16, cutoff=1e-3, chunk_mode="exact")).backward_f64(dpix, doth)
    for rep in range(2):
        for hl in (out["_hitlist"], None):
            gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=16, cutoff=1e-3,
                            chunk_mode="exact", hitlist=hl)
            for key in ("dopac", "dsigma", "dsh", "dverts"):
                ok, worst = grad_close(gb[key], rb[key])
                assert ok, (rep, hl is None, key, worst)
