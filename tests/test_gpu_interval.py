"""Depth intervals through the C ABI (40k splats, 144x112 camera, exact chunk mode):

  interval_out  per ray (t of the first / last blended hit, t of the terminating
                hit, first blended triangle) ......... bit-exact vs the oracle's decision log
  interval_in   (start, cap): any cap leaves the result unchanged (bit-exact vs
                the plain trace); a start equals the oracle trace started there
                (RGB <= 1e-6 abs vs oracle FP32, statistics identical)
  backward      bounded by the forward's interval_out: gradients equal the full
                re-traversal (<= 3e-4 relative; FP32 atomics sum in another order)
                and the FP64 oracle (<= 1e-3 relative)
"""
import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes
from helpers import GpuScene, grad_close, log_lists, oracle_scene

pytestmark = pytest.mark.gpu
K, CUT = 16, 1e-3


def _say(*a):
    print("[gpu-interval]", *a, flush=True)


@pytest.fixture(scope="module")
def small(cuda):
    # dense enough that about half of the rays terminate (cutoff 1e-3) over several chunks
    sc = scenes.uniform_scene(40_000, 6, offset_std=0.08)
    ro, rd = scenes.look_at_rays(144, 112, scenes.orbit_eye(3.0, 6), fov_deg=50.0)
    g = GpuScene(sc, cuda)
    S = oracle_scene(sc)
    _say("plain exact forward with interval_out")
    plain = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_out=True)
    log = O.Log(S, ro, rd, O.make_opts(k=K, cutoff=CUT, chunk_mode="exact"))
    return sc, ro, rd, g, S, plain, log


def _expected_intervals(log, R):
    exp = np.empty((R, 4), np.float32)
    exp[:, :3] = np.inf
    exp[:, 3] = np.uint32(0xFFFFFFFF).view(np.float32)
    for r, (p, t, c) in enumerate(log_lists(log.export(), R)):
        bl = np.nonzero(c == 2)[0]
        if bl.size:
            exp[r, 0], exp[r, 1] = t[bl[0]], t[bl[-1]]
            exp[r, 3] = np.uint32(p[bl[0]]).view(np.float32)
        te = np.nonzero(c == 3)[0]
        if te.size:
            exp[r, 2] = t[te[0]]
    return exp


def test_interval_out_matches_decision_log(small):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    exp = _expected_intervals(log, R)
    got = plain["interval"]
    bad = np.nonzero(~np.all(got.view(np.uint32) == exp.view(np.uint32), axis=1))[0]
    _say(f"rays with blended hits {np.isfinite(exp[:, 0]).sum()}, terminated {np.isfinite(exp[:, 2]).sum()}, "
         f"mismatching {bad.size}")
    assert bad.size == 0


def test_cap_leaves_the_result_unchanged(small):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    iv = plain["interval"]
    rng = np.random.Generator(np.random.PCG64(4))
    caps = {
        "random": rng.uniform(0.0, 6.0, R),
        "zero": np.zeros(R),
        "first": np.where(np.isfinite(iv[:, 0]), iv[:, 0], 3.0),
        "term": np.where(np.isfinite(iv[:, 2]), iv[:, 2], 3.0),
        "beyond far": np.full(R, 5000.0),
    }
    for name, cap in caps.items():
        _say("cap:", name)
        ivi = np.stack([np.zeros(R), cap], 1).astype(np.float32)
        out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=ivi, interval_out=True)
        for key in ("color", "others", "stats", "max_blending", "density", "interval"):
            # bit patterns: the triangle-id field of a ray without blended hits is 0xFFFFFFFF (a NaN)
            assert np.array_equal(out[key].view(np.uint32), plain[key].view(np.uint32)), (name, key)


@pytest.mark.parametrize("k", [8, 32])
def test_start_equals_oracle_started_there(small, k):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    iv = plain["interval"]
    rng = np.random.Generator(np.random.PCG64(5 + k))
    start = rng.uniform(0.0, 4.0, R).astype(np.float32)
    # a third of the rays start exactly at their first blended hit (inclusive key)
    sel = (rng.uniform(size=R) < 0.33) & np.isfinite(iv[:, 0])
    start[sel] = iv[sel, 0]
    cap = rng.uniform(0.0, 5.0, R).astype(np.float32)  # some caps lie before their start
    term = np.isfinite(iv[:, 2])
    assert 0.3 < term.mean() < 0.9
    _say(f"per-ray start, k={k}")
    out = g.forward(ro, rd, k=k, cutoff=CUT, chunk_mode="exact", interval_in=np.stack([start, cap], 1))
    ref = O.forward_f32(S, ro, rd, O.make_opts(k=k, cutoff=CUT, chunk_mode="exact", tstart=start))
    d = np.abs(out["color"] - ref["color"]).max()
    _say(f"  max |gpu-f32| {d:.3g}")
    assert d <= 1e-6
    assert np.abs(out["others"] - ref["others"]).max() <= 1e-6
    assert np.array_equal(out["stats"], ref["stats"])
    assert np.array_equal(out["color"][:, sel], plain["color"][:, sel])


def test_hard_cap_with_the_emitted_interval_reproduces_plain(small):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    iv = plain["interval"]
    has = np.isfinite(iv[:, 0])
    stop = np.where(np.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    # rays without blended hits: an empty interval past everything (the probe finds no hit -> full trace)
    ivi = np.stack([np.where(has, iv[:, 0], 900.0), np.where(has, stop, 901.0)], 1).astype(np.float32)
    _say("hard cap on the exact blended range")
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=ivi, hard_cap=True)
    assert np.array_equal(out["color"], plain["color"]) and np.array_equal(out["others"], plain["others"])
    # a start fallback wider than the scene sends every ray back to the full trace
    rng = np.random.Generator(np.random.PCG64(12))
    bad = np.stack([rng.uniform(0.0, 6.0, R), rng.uniform(0.0, 6.0, R)], 1).astype(np.float32)
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=bad, hard_cap=True,
                    fallback=(1000.0, 0.0))
    full = bad[:, 0] > 0.0
    assert np.array_equal(out["color"][:, full], plain["color"][:, full])
    # a hard cap inside the blended range of an open ray drops the rest of it; a ray
    # with no blended hit up to its cap goes on past it
    mid = 0.5 * (iv[:, 0] + iv[:, 1])
    open_ = has & ~np.isfinite(iv[:, 2]) & (iv[:, 1] > mid) & (mid > iv[:, 0])
    ivc = np.stack([np.zeros(R), np.where(has, mid, 0.3)], 1).astype(np.float32)
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=ivc, hard_cap=True)
    _say(f"open rays cut inside their blended range: {open_.sum()}")
    assert open_.sum() > 50
    assert np.all(out["others"][1, open_] < plain["others"][1, open_])
    assert np.array_equal(out["color"][:, ~has], plain["color"][:, ~has])


def test_hard_cap_with_perfect_intervals_reproduces_plain(small):
    # the proposal's ceiling: each ray's own blended range as (start, cap), rays that only
    # terminate start at their terminating hit, rays that blend nothing start past the far
    # plane (empty trace, no fallback); the image, normals and alpha equal plain bit for bit
    sc, ro, rd, g, S, plain, log = small
    iv = plain["interval"]
    first = np.where(np.isfinite(iv[:, 0]), iv[:, 0], iv[:, 2])
    stop = np.where(np.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    ivi = np.stack([np.where(np.isfinite(first), first, 3.0e38), np.where(np.isfinite(stop), stop, 0.0)],
                   1).astype(np.float32)
    empty = ~np.isfinite(first)
    _say(f"hard cap on perfect intervals; {int(empty.sum())} rays skipped as empty")
    assert empty.sum() > 50
    out = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=ivi, hard_cap=True)
    assert np.array_equal(out["color"], plain["color"]) and np.array_equal(out["others"], plain["others"])
    assert np.all(out["stats"][empty] == 0)
    soft = g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="exact", interval_in=ivi)
    assert np.array_equal(soft["color"], plain["color"])


def test_backward_bounded_by_forward_interval(small):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    rng = np.random.Generator(np.random.PCG64(9))
    dpix = rng.standard_normal((3, R)).astype(np.float32)
    doth = np.zeros((7, R), np.float32)
    doth[2:5] = 0.3 * rng.standard_normal((3, R)).astype(np.float32)
    _say("backward: full re-traversal vs bounded by interval_out")
    full = g.backward(ro, rd, plain["_color"], plain["_others"], dpix, doth, k=K, cutoff=CUT, chunk_mode="exact")
    bounded = g.backward(ro, rd, plain["_color"], plain["_others"], dpix, doth, k=K, cutoff=CUT, chunk_mode="exact",
                         interval=plain["interval"])
    ref = log.backward_f64(dpix, doth)
    for key in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(bounded[key], full[key], rel=3e-4)
        assert ok, (key, worst)
        ok, worst = grad_close(bounded[key], ref[key])
        _say(f"  {key}: worst ratio vs FP64 {worst:.3g}")
        assert ok, key


def test_interval_needs_exact_mode(small):
    sc, ro, rd, g, S, plain, log = small
    R = ro.shape[0] * ro.shape[1]
    ivi = np.zeros((R, 2), np.float32)
    with pytest.raises(RuntimeError, match="TT_CHUNK_EXACT"):
        g.forward(ro, rd, k=K, cutoff=CUT, chunk_mode="compat", interval_in=ivi)
