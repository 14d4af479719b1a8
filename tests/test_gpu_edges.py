"""Edge cases on the GPU against the oracle: images whose sides are not multiples of the
16x8 tile (partial tiles in the probe order, the per-tile record regions and the hit lists),
a 1x1 image, one triangle (the two-leaf tree), no triangles, and rays that miss everything.

Bars: colour / others within 1e-6 of the oracle's FP32 path, statistics identical, gradients
elementwise within 1e-3 max(|ref|, 1e-3 max|ref|) of the FP64 oracle; with no triangles or
no hits the colour is the background and the gradients are zero."""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes
from diff_tritracer.scenes import SplatScene
from helpers import GpuScene, grad_close, oracle_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-edges]", *a, flush=True)


def _check(g, S, ro, rd, mode, cap, bg=(0.0, 0.0, 0.0)):
    R = ro.shape[0] * ro.shape[1]
    out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode=mode, hitlist_cap=cap, bg=bg)
    opts = O.make_opts(k=16, cutoff=1e-3, chunk_mode=mode, bg=bg)
    ref = O.forward_f32(S, ro, rd, opts)
    assert np.abs(out["color"] - ref["color"]).max() <= 1e-6
    assert np.abs(out["others"] - ref["others"]).max() <= 1e-6
    assert np.array_equal(out["stats"], ref["stats"])
    rng = np.random.Generator(np.random.PCG64(21))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    doth[2:5] = 0.3 * rng.standard_normal((3, R)).astype(np.float32)
    gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=16, cutoff=1e-3, chunk_mode=mode,
                    hitlist=out.get("_hitlist"))
    rb = O.Log(S, ro, rd, opts).backward_f64(dpix, doth)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(gb[key], rb[key])
        assert ok, (key, worst)
    return out


@pytest.fixture(scope="module")
def uni(cuda):
    sc = scenes.uniform_scene(5_000, 9, offset_std=0.05)
    return sc, GpuScene(sc, cuda), oracle_scene(sc)


@pytest.mark.parametrize("W,H", [(37, 23), (1, 1), (16, 9), (129, 7)])
@pytest.mark.parametrize("mode,cap", [("compat", 0), ("exact", 4), ("exact", 64)])
def test_partial_tiles(uni, W, H, mode, cap):
    sc, g, S = uni
    _say(f"{W}x{H}, {mode}, hit list capacity {cap}")
    ro, rd = scenes.look_at_rays(W, H, scenes.orbit_eye(2.5, 9), fov_deg=40.0)
    _check(g, S, ro, rd, mode, cap)


def _one_triangle():
    v = np.array([[-0.5, -0.4, 0.0], [0.6, -0.3, 0.1], [0.0, 0.7, -0.1]], np.float32)
    rng = np.random.Generator(np.random.PCG64(5))
    shs = (0.2 * rng.standard_normal((1, 16, 3))).astype(np.float32)
    return SplatScene(verts=v, npts=np.array([3], np.int32), cumsum=np.array([0], np.int32), shs=shs,
                      opac=np.array([[0.8]], np.float32), sigma=np.array([[1.3]], np.float32))


@pytest.mark.parametrize("cap", [0, 8])
def test_one_triangle(cuda, cap):
    sc = _one_triangle()
    g, S = GpuScene(sc, cuda), oracle_scene(sc)
    ro, rd = scenes.look_at_rays(33, 17, (0.2, 0.1, 2.0), fov_deg=45.0)
    _say(f"one triangle, hit list capacity {cap}")
    out = _check(g, S, ro, rd, "exact", cap, bg=(0.1, 0.2, 0.3))
    assert (out["stats"][:, 1] > 0).any()


def test_no_triangles_and_missing_rays(cuda, uni):
    import torch

    empty = SplatScene(verts=np.zeros((0, 3), np.float32), npts=np.zeros(0, np.int32),
                       cumsum=np.zeros(0, np.int32), shs=np.zeros((0, 16, 3), np.float32),
                       opac=np.zeros((0, 1), np.float32), sigma=np.zeros((0, 1), np.float32))
    bg = (0.25, 0.5, 0.75)
    ro, rd = scenes.look_at_rays(20, 11, (0.0, 0.0, 3.0), fov_deg=30.0)
    _say("no triangles")
    g0 = GpuScene(empty, cuda)
    out = g0.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", bg=bg, hitlist_cap=8)
    assert np.array_equal(out["color"], np.repeat(np.array(bg, np.float32)[:, None], 20 * 11, 1))
    assert np.array_equal(out["stats"], np.zeros_like(out["stats"]))
    assert np.array_equal(out["hitlist"][1], np.zeros_like(out["hitlist"][1]))
    _say("rays that miss everything")
    sc, g, S = uni
    ro2, rd2 = scenes.look_at_rays(24, 10, (0.0, 0.0, 6.0), target=(0.0, 0.0, 12.0), fov_deg=20.0)
    out = g.forward(ro2, rd2, k=16, cutoff=1e-3, chunk_mode="exact", bg=bg, hitlist_cap=8)
    assert np.array_equal(out["color"], np.repeat(np.array(bg, np.float32)[:, None], 240, 1))
    R = 240
    gb = g.backward(ro2, rd2, out["_color"], out["_others"], np.ones((3, R), np.float32), np.zeros((7, R), np.float32),
                    k=16, cutoff=1e-3, chunk_mode="exact", hitlist=out["_hitlist"])
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        assert not np.any(gb[key]), key
    torch.cuda.synchronize()
