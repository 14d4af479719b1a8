"""Pins of the oracle's bounce variant (per-triangle reflectivity / IOR, one reflected and
one refracted ray from the primary's expected surface, equirect environment light; a
project-defined model with no counterpart in the reference), CPU only:

  environment lookup ...... closed form on textures linear in the texel column / row
  mirror and refraction ... closed forms for one triangle seen through its incenter
                            (alpha = opacity), normal and oblique incidence (Snell)
  gradients (FP64) ........ central finite differences of the replayed loss for the
                            reflectivity, the environment texels, opacity, sigma and SH
"""
import math

import numpy as np

import oracle as O
from diff_tritracer import scenes

C0 = 0.28209479177387814


def _linear_env(W, H, axis):
    e = np.zeros((H, W, 3), np.float32)
    if axis == "x":
        e[:, :, 0] = np.arange(W)[None, :]
    else:
        e[:, :, 0] = np.arange(H)[:, None]
    e[:, :, 1] = 1.0
    return e


def _env_closed_form(d, W, H):
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    u = (math.atan2(d[0], -d[2]) / (2 * math.pi) + 0.5) * W - 0.5
    v = math.acos(d[1]) / math.pi * H - 0.5
    return u, v


def test_environment_lookup_closed_form():
    W, H = 64, 32
    rng = np.random.Generator(np.random.PCG64(0))
    for _ in range(50):
        d = rng.normal(size=3).astype(np.float32)
        u, v = _env_closed_form(d, W, H)
        idx, w = O.env_taps(d, W, H)
        assert abs(w.sum() - 1.0) < 1e-12 and np.all(w >= 0)
        if 0.0 <= u <= W - 1:  # away from the wrap seam the lookup is linear in the column
            ex = _linear_env(W, H, "x").reshape(-1, 3)
            assert abs((w * ex[idx, 0]).sum() - u) < 1e-9
        ey = _linear_env(W, H, "y").reshape(-1, 3)
        assert abs((w * ey[idx, 0]).sum() - min(max(v, 0.0), H - 1)) < 1e-9


def _one_triangle(opac, sh0, refl, ior, z=2.0):
    # equilateral triangle in the plane z, incenter at the origin of x and y
    r = 0.5
    pts = np.array([[r * math.cos(a), r * math.sin(a), z] for a in (0.3, 0.3 + 2 * math.pi / 3, 0.3 + 4 * math.pi / 3)],
                   np.float32)
    shs = np.zeros((1, 16, 3), np.float32)
    shs[0, 0, :] = sh0
    S = O.Scene(pts, np.array([3]), np.array([0]), shs, 3, np.array([opac], np.float32),
                np.array([1.0], np.float32))
    return S, np.array([refl], np.float32), np.array([ior], np.float32)


def _render(S, refl, ior, env, o, d):
    opts = O.make_opts(k=16, cutoff=1e-4, chunk_mode="exact", variant="bounce", refl=refl, ior=ior, env=env)
    f = O.forward_f32(S, np.asarray([o], np.float32), np.asarray([d], np.float32), opts)
    lg = O.Log(S, np.asarray([o], np.float32), np.asarray([d], np.float32), opts)
    c64, _ = lg.forward_f64_bounce(refl, env)
    return f["color"][:, 0], c64[:, 0]


def _E(env, d):
    idx, w = O.env_taps(np.asarray(d, np.float32), env.shape[1], env.shape[0])
    return (w[:, None] * env.reshape(-1, 3)[idx]).sum(0)


def test_mirror_at_normal_incidence():
    rng = np.random.Generator(np.random.PCG64(1))
    env = np.exp(rng.normal(0, 0.5, (32, 64, 3))).astype(np.float32)
    sh0 = np.array([0.3, -0.2, 0.1])
    S, refl, ior = _one_triangle(0.8, sh0, 1.0, 1.3)
    c32, c64 = _render(S, refl, ior, env, (0, 0, 0), (0, 0, 1))
    a = float(np.float32(0.8))
    col = np.maximum(C0 * sh0.astype(np.float32).astype(np.float64) + 0.5, 0.0)
    # reflected ray straight back (nothing hit), transmitted part of the primary to E(+z)
    expect = a * col + a * _E(env, (0, 0, -1)) + (1 - a) * _E(env, (0, 0, 1))
    # float32 vertices put the incenter within ~1e-8 of the ray: Cx = 1 to that level
    np.testing.assert_allclose(c64, expect, rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(c32, expect, rtol=2e-6, atol=2e-6)


def test_refraction_follows_snell():
    rng = np.random.Generator(np.random.PCG64(2))
    env = np.exp(rng.normal(0, 0.5, (64, 128, 3))).astype(np.float32)
    sh0 = np.array([0.1, 0.2, -0.3])
    eta = 1.45
    S, refl, ior = _one_triangle(0.6, sh0, 0.0, eta)
    th = 0.5  # incidence angle; the ray passes through the incenter (0, 0, 2)
    d = np.array([math.sin(th), 0.0, math.cos(th)])
    o = np.array([0.0, 0.0, 2.0]) - 2.0 * d
    c32, c64 = _render(S, refl, ior, env, o, d)
    a = float(np.float32(0.6))
    col = np.maximum(C0 * sh0.astype(np.float32).astype(np.float64) + 0.5, 0.0)
    st = math.sin(th) / eta
    dt = np.array([st, 0.0, math.sqrt(1 - st * st)])
    expect = a * col + a * _E(env, dt) + (1 - a) * _E(env, d)
    np.testing.assert_allclose(c64, expect, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(c32, expect, rtol=1e-5, atol=1e-5)


def test_bounce_gradients_match_finite_differences():
    sc = scenes.uniform_scene(3000, 6, offset_std=0.08)
    ro, rd = scenes.look_at_rays(24, 20, scenes.orbit_eye(3.0, 6), fov_deg=50.0)
    rng = np.random.Generator(np.random.PCG64(1))
    refl = rng.beta(2, 5, sc.P).astype(np.float32)
    ior = rng.uniform(1.0, 1.6, sc.P).astype(np.float32)
    env = np.exp(rng.normal(0, 0.5, (16, 32, 3))).astype(np.float32)
    S = O.Scene(sc.verts, sc.npts, sc.cumsum, sc.shs, 3, sc.opac, sc.sigma)
    opts = O.make_opts(k=16, cutoff=1e-3, chunk_mode="exact", variant="bounce", refl=refl, ior=ior, env=env)
    f = O.forward_f32(S, ro, rd, opts)
    lg = O.Log(S, ro, rd, opts)
    c64, o64 = lg.forward_f64_bounce(refl, env)
    assert np.abs(f["color"] - c64).max() < 1e-5
    R = ro.shape[0] * ro.shape[1]
    dp = rng.standard_normal((3, R))
    do = np.zeros((7, R))
    do[2:5] = 0.3 * rng.standard_normal((3, R))
    g = lg.backward_f64_bounce(refl, env, dp, do)
    base = dict(refl=refl.astype(np.float64), env=env.astype(np.float64), opac=sc.opac.reshape(-1).astype(np.float64),
                sigma=sc.sigma.reshape(-1).astype(np.float64), shs=sc.shs.astype(np.float64))
    gkey = dict(refl="drefl", env="denv", opac="dopac", sigma="dsigma", shs="dsh")

    def loss(name, val):
        kw = {name: val}
        r = kw.pop("refl", base["refl"])
        e = kw.pop("env", base["env"])
        return lg.loss_f64_bounce(r, e, dp, do, **kw)

    h = 1e-6
    for name in gkey:
        ga = g[gkey[name]].reshape(-1)
        assert np.count_nonzero(ga) > 50, name
        for i in np.argsort(-np.abs(ga))[:12]:
            p1 = base[name].reshape(-1).copy()
            p2 = p1.copy()
            p1[i] += h
            p2[i] -= h
            fd = (loss(name, p1.reshape(base[name].shape)) - loss(name, p2.reshape(base[name].shape))) / (2 * h)
            assert abs(fd - ga[i]) <= 1e-6 * abs(fd) + 1e-9, (name, i, fd, ga[i])
