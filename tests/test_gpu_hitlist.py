"""Stored blended hits on the GPU: the forward writes each pixel's blended hits (triangle id,
t) and their count; the backward reads them instead of re-traversing the BVH (small config:
10k splats, seed 0, 256x256 camera, k = 16, cutoff 1e-3, both chunk modes).

Bars:
  forward outputs with the hit list ...... bit-identical to the forward without it
  stored hits ............................ (triangle id, t bits) equal to the oracle's blended hits of
                                           the pixel, in order, for i < min(count, capacity)
  counts ................................. the oracle's blended count; -1 where it exceeds the capacity
  backward from the list ................. elementwise within 1e-3 max(|ref|, 1e-3 max|ref|) of the FP64
                                           oracle, also at a capacity of 3 (pixels with more blended hits
                                           are re-traversed: a mixed list / re-traversal launch)
  stored colours ......................... every stored hit's colour >= 0 or -0.0 (a clamped channel);
                                           the backward that reads them equals, within 1e-3 (FP32 atomics
                                           add in any order; a wrong clamp bit errs by O(1)), the backward
                                           that evaluates the SH again
  window records reused ................. a backward told the scene is unchanged reuses the forward's
                                           per-triangle window records only for the same input pointers:
                                           with other vertices it forms their records (same gradients as
                                           without the flag); with the forward's own, the same gradients
                                           as forming them again (within 1e-3: FP32 atomic order)
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes
from helpers import GpuScene, grad_close, log_lists, oracle_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-hitlist]", *a, flush=True)


@pytest.fixture(scope="module")
def small(cuda):
    sc, ro, rd = scenes.small_config(0)
    return sc, ro, rd, GpuScene(sc, cuda), oracle_scene(sc)


@pytest.mark.parametrize("mode", ["compat", "exact"])
@pytest.mark.parametrize("cap", [64, 3])
def test_hitlist_forward_and_backward(small, mode, cap):
    sc, ro, rd, g, S = small
    R = ro.shape[0] * ro.shape[1]
    _say(f"forward with hit list, chunk mode {mode}, capacity {cap}")
    plain = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode=mode)
    out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode=mode, hitlist_cap=cap)
    assert np.array_equal(out["color"].view(np.uint32), plain["color"].view(np.uint32))
    assert np.array_equal(out["others"].view(np.uint32), plain["others"].view(np.uint32))
    opts = O.make_opts(k=16, cutoff=1e-3, chunk_mode=mode)
    log = O.Log(S, ro, rd, opts)
    lists = log_lists(log.export(), R)
    hits, counts = out["hitlist"][0][0], out["hitlist"][1][0]  # segment 0: the primary rays
    bad_count = bad_hits = overflow = 0
    for r in range(R):
        if lists[r] is None:
            bp, bt = np.zeros(0, np.int32), np.zeros(0, np.float32)
        else:
            p, t, c = lists[r]
            sel = c == 2
            bp, bt = p[sel], t[sel]
        n = len(bp)
        want = n if n <= cap else -1
        overflow += want < 0
        if counts[r] != want:
            bad_count += 1
            continue
        m = min(n, cap)
        if not (np.array_equal(hits[:m, r, 0].view(np.uint32), bp[:m].astype(np.uint32)) and
                np.array_equal(hits[:m, r, 1].view(np.uint32), bt[:m].view(np.uint32))):
            bad_hits += 1
    _say(f"  counts wrong {bad_count}, hit lists wrong {bad_hits}, pixels over capacity {overflow} of {R}")
    assert bad_count == 0 and bad_hits == 0
    rng = np.random.Generator(np.random.PCG64(11))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    doth[2:5] = 0.3 * rng.standard_normal((3, R)).astype(np.float32)
    _say("  backward from the hit list")
    gb = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=16, cutoff=1e-3, chunk_mode=mode,
                    hitlist=out["_hitlist"])
    rb = log.backward_f64(dpix, doth)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(gb[key], rb[key])
        _say(f"  {key}: worst ratio vs FP64 {worst:.3g}")
        assert ok, key
    cols = out["hitlist"][2][0]  # [cap, R, 3] of segment 0
    stored = np.zeros((cap, R), bool)
    for r in range(R):
        stored[: max(0, min(int(counts[r]) if counts[r] >= 0 else cap, cap)), r] = True
    c = cols[stored]
    clamped = (c == 0.0) & np.signbit(c)
    _say(f"  stored colours: {c.shape[0]} hits, {int(clamped.sum())} clamped channels")
    assert np.all((c > 0.0) | (c == 0.0)) and np.all(np.isfinite(c))
    _say("  backward from the hit list without its colours (SH evaluated again)")
    gn = g.backward(ro, rd, out["_color"], out["_others"], dpix, doth, k=16, cutoff=1e-3, chunk_mode=mode,
                    hitlist=out["_hitlist"][:2])
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(gb[key], gn[key])
        _say(f"  {key}: worst ratio vs the SH-evaluating backward {worst:.3g}")
        assert ok, (key, worst)


def test_backward_reuses_window_records_only_for_the_forward_inputs(small):
    import torch

    sc, ro, rd, g, S = small
    R = ro.shape[0] * ro.shape[1]
    out = g.forward(ro, rd, k=16, cutoff=1e-3, chunk_mode="exact", hitlist_cap=64)
    rng = np.random.Generator(np.random.PCG64(12))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    args = (ro, rd, out["_color"], out["_others"], dpix, doth)
    kw = dict(k=16, cutoff=1e-3, chunk_mode="exact", hitlist=out["_hitlist"])
    _say("backward reusing the forward's window records, and forming them again")
    reuse = g.backward(*args, same_scene=True, **kw)
    formed = g.backward(*args, same_scene=False, **kw)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(reuse[key], formed[key])
        assert ok, (key, worst)
    # other vertices (a different tensor, moved by 1e-3) on the same tracer: records reused from
    # the forward would be those of the old vertices; the flag must not apply to them
    saved = g.verts
    try:
        g.verts = saved + 1e-3 * torch.randn(saved.shape, generator=torch.Generator().manual_seed(5)).to(saved.device)
        _say("backward with other vertices and the flag set, and without it")
        flagged = g.backward(*args, same_scene=True, **kw)
        plain = g.backward(*args, same_scene=False, **kw)
    finally:
        g.verts = saved
    moved = 0
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(flagged[key], plain[key])
        assert ok, (key, worst)
        moved += not grad_close(flagged[key], reuse[key])[0]
    assert moved > 0  # the moved vertices do change the gradients: the check above can tell
