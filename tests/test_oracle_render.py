"""Pins for the oracle's compositing, chunking, statistics, variants and gradients,
against closed forms, brute force and FP64 finite differences (CPU only).

Semantics pinned here (reference file:line):
  k-closest chunks resumed at last_hit_t, inclusive ....... closestk.h:17-149, forward.cu:99-111
  near-plane skip t < 0.2, consumes a slot ................ forward.cu:107-113, auxiliary.h:29
  alpha = min(0.99, opacity*Cx), skip below 1/255 ......... forward.cu:129-131
  stop before blending when T(1-alpha) < cutoff ........... forward.cu:133-136
  outputs: colour + T*bg, others = (0, 1-T, N, 0, 0) ...... forward.cu:203-214
  max_blending / density / scaling ........................ forward.cu:34-74, :138-146
  backward formulas ....................................... backward.cu:191-350
"""
import math

import numpy as np
import pytest

import oracle as O

C0 = 0.28209479177387814


def stacked_scene(zs, opac, sh0, sigma=1.3, shape=((0.0, 0.0), (1.0, 0.0), (0.25, 0.9))):
    """Identical triangles in the planes z = zs[i]; returns arrays and the incenter (x, y)."""
    P = len(zs)
    base = np.array([[x, y, 0.0] for x, y in shape])
    verts = np.concatenate([base + np.array([0, 0, z]) for z in zs]).astype(np.float32)
    a = np.linalg.norm(base[1] - base[2])
    b = np.linalg.norm(base[0] - base[2])
    c = np.linalg.norm(base[0] - base[1])
    inc = (a * base[0] + b * base[1] + c * base[2]) / (a + b + c)
    shs = np.zeros((P, 16, 3), np.float32)
    shs[:, 0, :] = np.asarray(sh0, np.float32).reshape(P, 3)
    S = O.Scene(verts, np.full(P, 3), 3 * np.arange(P), shs, 3, np.asarray(opac, np.float32),
                np.full(P, sigma, np.float32))
    return S, inc


def closed_form(seq, opac, sh0, cutoff, bg=(0.0, 0.0, 0.0)):
    """Front-to-back compositing of the hit sequence (triangle indices), in double."""
    C = np.zeros(3)
    T = 1.0
    n = 0
    for i in seq:
        a = min(0.99, opac[i])
        if a < 1 / 255:
            continue
        if T * (1 - a) < cutoff:
            break
        col = np.maximum(C0 * np.asarray(sh0[i], float) + 0.5, 0.0)
        C += col * a * T
        T *= 1 - a
        n += 1
    return C + T * np.asarray(bg), T, n


def _render(S, inc, k, mode, cutoff=1e-3, bg=(0, 0, 0)):
    o = np.array([[inc[0], inc[1], 0.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    opts = O.make_opts(k=k, cutoff=cutoff, chunk_mode=mode, bg=bg)
    out = O.forward_f32(S, o, d, opts)
    c64, o64 = O.Log(S, o, d, opts).forward_f64()
    return out, c64, o64


def test_stacked_exact_mode_is_plain_front_to_back():
    P = 40
    zs = 1.0 + 0.1 * np.arange(P)
    rng = np.random.Generator(np.random.PCG64(0))
    opac = rng.uniform(0.02, 0.08, P)
    sh0 = rng.normal(0, 0.5, (P, 3))
    S, inc = stacked_scene(zs, opac, sh0)
    for k in (8, 16, 32):
        out, c64, _ = _render(S, inc, k, "exact")
        C, T, n = closed_form(range(P), opac, sh0, 1e-3)
        np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
        np.testing.assert_allclose(c64[:, 0], C, atol=1e-9)
        assert abs(out["others"][1, 0] - (1 - T)) < 1e-6
        assert out["stats"][0, 0] == P and out["stats"][0, 1] == n


def test_stacked_start_distance_drops_the_planes_before_it():
    # a per-ray start distance (exact mode) composites only the planes at t >= start
    P = 40
    zs = 1.0 + 0.1 * np.arange(P)
    rng = np.random.Generator(np.random.PCG64(2))
    opac = rng.uniform(0.02, 0.08, P)
    sh0 = rng.normal(0, 0.5, (P, 3))
    S, inc = stacked_scene(zs, opac, sh0)
    o = np.array([[inc[0], inc[1], 0.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    prims, ts = O.gather(S, o[0], d[0], k=P)
    assert list(prims) == list(range(P))
    t5 = float(ts[5])  # the start key is inclusive: t == start is kept, the next float is not
    cases = ((0, 0.0), (0, 0.9), (5, 1.45), (5, t5), (6, float(np.nextafter(np.float32(t5), np.float32(9)))),
             (17, 2.65), (39, 4.85), (40, 6.0))
    for first, start in cases:
        for k in (8, 16):
            opts = O.make_opts(k=k, cutoff=1e-3, chunk_mode="exact", tstart=[start])
            out = O.forward_f32(S, o, d, opts)
            c64, _ = O.Log(S, o, d, opts).forward_f64()
            C, T, n = closed_form(range(first, P), opac, sh0, 1e-3)
            np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
            np.testing.assert_allclose(c64[:, 0], C, atol=1e-9)
            assert out["stats"][0, 0] == P - first and out["stats"][0, 1] == n


def test_stacked_compat_mode_repeats_chunk_boundary_hit():
    # resuming at t >= last_hit_t re-reports the last hit of every full chunk
    P = 40
    zs = 1.0 + 0.1 * np.arange(P)
    rng = np.random.Generator(np.random.PCG64(1))
    opac = rng.uniform(0.02, 0.08, P)
    sh0 = rng.normal(0, 0.5, (P, 3))
    S, inc = stacked_scene(zs, opac, sh0)
    out, c64, _ = _render(S, inc, 16, "compat")
    seq = list(range(16)) + list(range(15, 31)) + list(range(30, 40))
    C, T, n = closed_form(seq, opac, sh0, 1e-3)
    np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
    np.testing.assert_allclose(c64[:, 0], C, atol=1e-9)
    assert out["stats"][0, 0] == len(seq)


def test_compat_mode_terminates_on_equal_t_runs():
    # 20 coplanar copies at one depth, k = 8. Chunk 1 takes copies 0-7; the inclusive
    # resume at t = 2 returns the same 8 again (what the reference does, and would keep
    # doing forever); with no progress in t the tracer then resumes after the last
    # triangle index: 8-15, then 16-19.
    P = 20
    S, inc = stacked_scene(np.full(P, 2.0), np.full(P, 0.001), np.zeros((P, 3)))
    out, _, _ = _render(S, inc, 8, "compat")
    assert out["stats"][0, 0] == 8 + 8 + 8 + 4
    out, _, _ = _render(S, inc, 8, "exact")
    assert out["stats"][0, 0] == P


def test_near_plane_skip_consumes_slots():
    zs = [0.05, 0.1, 0.15, 0.3, 0.4]
    opac = [0.5] * 5
    sh0 = np.full((5, 3), 0.3)
    S, inc = stacked_scene(zs, opac, sh0)
    out, _, _ = _render(S, inc, 2, "compat")
    # chunks of 2 with inclusive resume: (.05 .1) skip skip | (.1 .15) skip skip |
    # (.15 .3) skip blend | (.3 .4) blend blend | (.4) blend -> blended sequence 3,3,4,4
    seq = [3, 3, 4, 4]
    C, T, n = closed_form(seq, opac, sh0, 1e-3)
    np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
    assert out["stats"][0, 0] == 9 and out["stats"][0, 1] == 4
    # a pair of coplanar copies, then one more copy; k = 2. At t = 0.01 the 1e-9 step of
    # forward.cu:111 lifts last_hit_t above the tied pair (ulp(0.01f)/2 < 1e-9), so the pair
    # is not re-reported: (A B) | (C) -> 3 consumed. At t = 2 the step vanishes in FP32:
    # (A B) | (A B) again | (C) -> 5 consumed.
    S2, inc2 = stacked_scene([0.01, 0.01, 0.3], [0.5] * 3, np.full((3, 3), 0.3))
    out2, _, _ = _render(S2, inc2, 2, "compat")
    assert out2["stats"][0, 0] == 3 and out2["stats"][0, 1] == 1
    S3, inc3 = stacked_scene([2.0, 2.0, 2.3], [0.5] * 3, np.full((3, 3), 0.3))
    out3, _, _ = _render(S3, inc3, 2, "compat")
    assert out3["stats"][0, 0] == 5 and out3["stats"][0, 1] == 5


def test_alpha_threshold_and_termination():
    zs = [1.0, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6]
    opac = [0.003, 0.8, 0.8, 0.8, 0.8, 0.8, 0.8]  # first below 1/255
    sh0 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1], [1, 1, 1]], float)
    S, inc = stacked_scene(zs, opac, sh0)
    out, c64, _ = _render(S, inc, 16, "exact", cutoff=1e-3)
    C, T, n = closed_form(range(7), opac, sh0, 1e-3)
    assert n == 4  # T: 1 -> .2 -> .04 -> .008 -> .0016, the 5th would leave 3.2e-4 < 1e-3
    np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
    assert abs(out["others"][1, 0] - (1 - T)) < 1e-6
    assert out["stats"][0, 0] == 6  # skip + 4 blends + the terminating hit; the last never reached
    # alpha clamp at 0.99
    S2, inc2 = stacked_scene([1.0], [1.0], [[0.2, 0.2, 0.2]])
    out2, _, _ = _render(S2, inc2, 16, "exact", cutoff=1e-3)
    assert abs(out2["others"][1, 0] - 0.99) < 1e-6


def test_outputs_normal_depth_and_statistics():
    zs = [1.0, 1.5, 2.0]
    opac = [0.3, 0.4, 0.5]
    sh0 = np.full((3, 3), 0.1)
    S, inc = stacked_scene(zs, opac, sh0)
    bg = (0.2, 0.4, 0.6)
    out, c64, o64 = _render(S, inc, 16, "exact", bg=bg)
    C, T, n = closed_form(range(3), opac, sh0, 1e-3, bg)
    np.testing.assert_allclose(out["color"][:, 0], C, atol=2e-6)
    # face normal of the (planar, z-stacked) triangles is +z; N = sum of weights = 1 - T
    np.testing.assert_allclose(out["others"][2:5, 0], [0, 0, 1 - T], atol=2e-6)
    assert out["others"][0, 0] == 0 and out["others"][5, 0] == 0 and out["others"][6, 0] == 0
    w = []
    Tr = 1.0
    for a in opac:
        w.append(a * Tr)
        Tr *= 1 - a
    np.testing.assert_allclose(out["max_blending"], w, rtol=1e-6)
    assert np.all(out["density"] == 1.0)
    # scaling: max angle between (vertex - origin) and (centroid - origin)
    o = np.array([inc[0], inc[1], 0.0])
    for i in range(3):
        V = S.verts[3 * i:3 * i + 3].astype(float)
        c = V.mean(0) - o
        ang = max(math.acos(np.clip(np.dot(v - o, c) / np.linalg.norm(v - o) / np.linalg.norm(c), -1, 1)) for v in V)
        assert abs(out["scaling"][i] - ang) < 2e-6


# ------------------------------------------------------------ BVH vs brute force
def _degenerate_scene():
    """A coplanar grid of triangles sharing edges (many equal t), plus axis-aligned
    slivers and random splats."""
    tris = []
    for i in range(12):
        for j in range(12):
            x0, y0 = -0.6 + 0.1 * i, -0.6 + 0.1 * j
            tris.append([[x0, y0, 0.5], [x0 + 0.1, y0, 0.5], [x0, y0 + 0.1, 0.5]])
            tris.append([[x0 + 0.1, y0, 0.5], [x0 + 0.1, y0 + 0.1, 0.5], [x0, y0 + 0.1, 0.5]])
    for i in range(40):
        x = -0.5 + 0.025 * i
        tris.append([[x, -0.5, 0.2], [x, 0.5, 0.2], [x, 0.0, 0.9]])  # planes x = const
    rng = np.random.Generator(np.random.PCG64(4))
    c = rng.uniform(-0.7, 0.7, (300, 1, 3))
    tris += list(c + 0.05 * rng.standard_normal((300, 3, 3)))
    V = np.asarray(tris, np.float32).reshape(-1, 3)
    P = V.shape[0] // 3
    rng2 = np.random.Generator(np.random.PCG64(5))
    return O.Scene(V, np.full(P, 3), 3 * np.arange(P), rng2.normal(0, 0.3, (P, 16, 3)), 3,
                   rng2.uniform(0.05, 0.9, P), np.exp(rng2.normal(0, 0.5, P)))


def test_bvh_gather_equals_brute_force():
    S = _degenerate_scene()
    rng = np.random.Generator(np.random.PCG64(6))
    for n in range(400):
        o = np.array([0.0, 0.0, -2.0], np.float32) + rng.normal(0, 0.05, 3).astype(np.float32)
        if n % 3 == 0:
            d = np.array([0, 0, 1], np.float32)           # axis-aligned rays
            o[:2] = np.round(o[:2] * 20) / 20              # through shared grid edges/vertices
        else:
            d = np.array([rng.normal(0, 0.3), rng.normal(0, 0.3), 1.0], np.float32)
        for tlo, plo in ((0.0, -1), (2.5, -1), (2.5, 100)):
            for k in (4, 16):
                pb, tb = O.gather(S, o, d, tlo, plo, k, use_bvh=True)
                pf, tf = O.gather(S, o, d, tlo, plo, k, use_bvh=False)
                assert np.array_equal(pb, pf) and np.array_equal(tb.view(np.uint32), tf.view(np.uint32))


def test_bvh_render_equals_brute_force():
    from diff_tritracer import scenes

    sc, ro, rd = scenes.small_config(0, 48, 48)
    S = O.Scene(sc.verts, sc.npts, sc.cumsum, sc.shs, 3, sc.opac, sc.sigma)
    a = O.forward_f32(S, ro, rd, O.make_opts(k=16, cutoff=1e-3, use_bvh=True))
    b = O.forward_f32(S, ro, rd, O.make_opts(k=16, cutoff=1e-3, use_bvh=False))
    assert np.array_equal(a["color"], b["color"]) and np.array_equal(a["stats"], b["stats"])


# ------------------------------------------------------------------ variants
def _single(z, opac=0.8, sh0=(0.5, 0.1, -0.2), size=0.6):
    V = np.array([[-size, -size, z], [size, -size, z], [0.0, size, z]], np.float32)
    shs = np.zeros((1, 16, 3), np.float32)
    shs[0, 0] = sh0
    return O.Scene(V, [3], [0], shs, 3, [opac], [0.5])


def test_reflect_sends_central_ray_straight_back():
    S = _single(7.0)                     # behind the camera
    o = np.array([[0.0, 0.0, 5.0]], np.float32)
    d = np.array([[0.0, 0.0, -1.0]], np.float32)
    sphere = [1.0, 0.0, 0.0, 0.0]
    plain = O.forward_f32(S, o, d, O.make_opts(cutoff=1e-4))
    refl = O.forward_f32(S, o, d, O.make_opts(cutoff=1e-4, variant="reflect", extra=sphere))
    assert plain["stats"][0, 0] == 0 and refl["stats"][0, 1] == 1
    miss = O.forward_f32(S, np.array([[3.0, 0, 5]], np.float32), d,
                         O.make_opts(cutoff=1e-4, variant="reflect", extra=sphere))
    assert miss["stats"][0, 0] == 0      # sphere missed: the primary ray is traced (and misses)


def test_refract_central_ray_continues_straight():
    S = _single(-3.0)
    o = np.array([[0.0, 0.0, 5.0]], np.float32)
    d = np.array([[0.0, 0.0, -1.0]], np.float32)
    log = O.Log(S, o, d, O.make_opts(cutoff=1e-4, variant="refract", extra=[1.0, 0, 0, 0]))
    off, spix, prims, ts, codes = log.export()
    # traced from the sphere's front point z = 1 (+1e-9 normal offset) along -z: t = 4
    assert len(ts) == 1 and abs(float(ts[0]) - 4.0) < 1e-6 and codes[0] == O.CODE_BLEND


def test_envlight_identities():
    S = _single(50.0, opac=0.0)  # nothing visible
    o = np.array([[0.0, 0.0, 5.0], [3.0, 0.0, 5.0]], np.float32)
    d = np.array([[0.0, 0.0, -1.0], [0.0, 0.0, -1.0]], np.float32)
    bg = (0.1, 0.2, 0.3)
    out = O.forward_f32(S, o, d, O.make_opts(cutoff=1e-4, variant="envlight", extra=[1.0, 0, 0, 0, 8], bg=bg))
    np.testing.assert_allclose(out["color"][:, 0], bg, atol=1e-7)  # mean of 8 bg samples, T = 0
    np.testing.assert_allclose(out["color"][:, 1], bg, atol=0)     # sphere missed: C + T*bg with T = 1


# ------------------------------------------------------ gradients: FP64 FD
def _fd_scene():
    rng = np.random.Generator(np.random.PCG64(9))
    P = 24
    c = rng.uniform(-0.3, 0.3, (P, 1, 3))
    V = (c + 0.15 * rng.standard_normal((P, 3, 3))).astype(np.float32).reshape(-1, 3)
    shs = rng.normal(0, 0.4, (P, 16, 3)).astype(np.float32)
    opac = rng.uniform(0.1, 0.85, P).astype(np.float32)      # below the 0.99 clamp
    sigma = np.exp(rng.normal(0, 0.3, P)).astype(np.float32)
    S = O.Scene(V, np.full(P, 3), 3 * np.arange(P), shs, 3, opac, sigma)
    from diff_tritracer import scenes
    ro, rd = scenes.look_at_rays(20, 20, (0.2, 0.3, 2.0), fov_deg=40)
    return S, ro.reshape(-1, 3), rd.reshape(-1, 3)


def test_f64_gradient_matches_finite_differences():
    S, ro, rd = _fd_scene()
    opts = O.make_opts(k=4, cutoff=1e-3)
    log = O.Log(S, ro, rd, opts)
    R = ro.shape[0]
    rng = np.random.Generator(np.random.PCG64(10))
    gp = rng.normal(0, 1, (3, R))
    go = np.zeros((7, R))
    go[2:5] = rng.normal(0, 1, (3, R))                       # normal channels
    g = log.backward_f64(gp, go)
    base = log.params64()
    names = ("verts", "opac", "sigma", "shs")
    gkeys = ("dverts", "dopac", "dsigma", "dsh")
    h = 1e-6
    for name, gk, arr in zip(names, gkeys, base):
        flat = arr.reshape(-1)
        an = g[gk].reshape(-1)
        scale = np.abs(an).max()
        idx = np.arange(flat.size) if flat.size <= 600 else rng.choice(flat.size, 600, replace=False)
        for i in idx:
            p = [b.copy() for b in base]
            pm = [b.copy() for b in base]
            p[names.index(name)].reshape(-1)[i] += h
            pm[names.index(name)].reshape(-1)[i] -= h
            lp = log.loss_f64(gp, go, **dict(zip(names, p)))
            lm = log.loss_f64(gp, go, **dict(zip(names, pm)))
            fd = (lp - lm) / (2 * h)
            assert abs(fd - an[i]) <= 1e-5 * max(abs(fd), 1e-2 * scale), (name, i, fd, an[i])


def test_alpha_channel_gradient_ignored_and_depth_term():
    # the reference backward reads dL/d(depth) but never dL/d(alpha channel)
    S, inc = stacked_scene([1.0], [0.6], [[0.3, 0.3, 0.3]])
    o = np.array([[inc[0], inc[1], 0.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    log = O.Log(S, o, d, O.make_opts(cutoff=1e-3))
    gp = np.zeros((3, 1))
    go = np.zeros((7, 1))
    go[1] = 5.0
    assert np.all(log.backward_f64(gp, go)["dopac"] == 0.0)
    # depth channel: with one hit, forward depth is 0 and D = alpha*t, so
    # dL/dalpha = dL/ddepth * (t + alpha*t/(1-alpha)) = dL/ddepth * t/(1-alpha); dL/dopac = that * Cx
    go[:] = 0
    go[0] = 1.0
    a = 0.6
    # (rtol: the FP32 opacity 0.6 and the FP32-rounded incenter ray, Cx = 1 - O(1e-8))
    np.testing.assert_allclose(log.backward_f64(gp, go)["dopac"], [1.0 * 1.0 / (1 - a)], rtol=1e-6)


def test_f32_backward_close_to_f64():
    from diff_tritracer import scenes

    sc, ro, rd = scenes.small_config(0, 96, 96)
    S = O.Scene(sc.verts, sc.npts, sc.cumsum, sc.shs, 3, sc.opac, sc.sigma)
    opts = O.make_opts(k=16, cutoff=1e-3)
    R = ro.shape[0] * ro.shape[1]
    f = O.forward_f32(S, ro, rd, opts)
    rng = np.random.Generator(np.random.PCG64(12))
    gp = rng.standard_normal((3, R)).astype(np.float32)
    go = np.zeros((7, R), np.float32)
    b32 = O.backward_f32(S, ro, rd, opts, f["color"], f["others"], gp, go)
    log = O.Log(S, ro, rd, opts)
    c64, _ = log.forward_f64()
    assert np.abs(f["color"] - c64).max() <= 1e-5
    b64 = log.backward_f64(gp, go)
    from helpers import grad_close
    for k in ("dopac", "dsigma", "dsh", "dverts"):
        ok, worst = grad_close(b32[k], b64[k])
        assert ok, (k, worst)


def test_vertex_abs_mode_bounds_the_signed_sum():
    # the test aid behind the bounce vertex-gradient bar: |sum of terms| <= sum of |terms|,
    # equal where one ray contributes one term; the signed gradient is unchanged afterwards
    S, ro, rd = _fd_scene()
    opts = O.make_opts(k=4, cutoff=1e-3)
    R = ro.shape[0]
    rng = np.random.Generator(np.random.PCG64(10))
    gp = rng.normal(0, 1, (3, R))
    go = np.zeros((7, R))
    log = O.Log(S, ro, rd, opts)
    signed = log.backward_f64(gp, go)["dverts"]
    absd = log.backward_f64(gp, go, vertex_abs=True)["dverts"]
    assert np.all(absd >= np.abs(signed) - 1e-15 * absd.max())
    assert np.any(absd > 2.0 * np.abs(signed))          # some entries do cancel
    assert np.array_equal(log.backward_f64(gp, go)["dverts"], signed)
    one = O.Log(S, ro[:1], rd[:1], opts)               # a single ray: every vertex entry one term or a few
    s1, a1 = one.backward_f64(gp[:, :1], go[:, :1])["dverts"], one.backward_f64(gp[:, :1], go[:, :1], True)["dverts"]
    assert np.all(a1 >= np.abs(s1) - 1e-15)
