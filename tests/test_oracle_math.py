"""Pins for the oracle's elementary pieces, each against something the reference
formula does not restate: textbook constants and orthonormality, barycentric
closed forms, exact rational arithmetic, numpy's libm.

The reference holds no golden vectors or tests (it has no tests/ directory and no fixtures),
so these closed forms and invariants are what fixes the oracle.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


# ------------------------------------------------------------ FP32 math kernels
def _ulp_err(got, ref):
    got = np.float32(got)
    ref32 = np.float32(ref)
    if not np.isfinite(ref32):
        return 0.0 if got == ref32 else np.inf
    spacing = np.spacing(np.abs(ref32)) if ref32 != 0 else np.float32(1e-45)
    return float(abs(np.float64(got) - ref) / spacing)


def test_log2_exp2_pow_accuracy():
    rng = np.random.Generator(np.random.PCG64(0))
    xs = np.concatenate([rng.uniform(1e-6, 1.0, 2000), rng.uniform(1.0, 1e6, 1000),
                         np.float32(2.0) ** rng.integers(-120, 120, 200)]).astype(np.float32)
    worst_log = max(_ulp_err(O.log2f(x), math.log2(float(x))) for x in xs if abs(math.log2(float(x))) > 1e-3)
    ys = rng.uniform(-125.9, 127.9, 3000).astype(np.float32)
    worst_exp = max(_ulp_err(O.exp2f(y), 2.0 ** float(y)) for y in ys)
    assert worst_log <= 4.0, worst_log
    assert worst_exp <= 4.0, worst_exp
    # the window function's use: x in (0, 1], y = sigma in (0.1, 10)
    px = rng.uniform(1e-4, 1.0, 2000).astype(np.float32)
    py = np.exp(rng.uniform(-2.3, 2.3, 2000)).astype(np.float32)
    # error model of exp2(y*log2 x) in FP32: the rounding of the exponent, eps*|y log2 x|,
    # times ln 2, plus a few ulp of the two kernels
    worst = 0.0
    for x, y in zip(px, py):
        ref = float(x) ** float(y)
        e = abs(float(O.powf(x, y)) - ref) / ref
        bound = 2 ** -24 * (abs(float(y) * math.log2(float(x))) * math.log(2) + 8)
        worst = max(worst, e / bound)
    assert worst <= 1.5, worst


def test_elementary_special_values():
    assert O.log2f(0.0) == -np.inf
    assert np.isnan(O.log2f(-1.0))
    assert O.log2f(1.0) == 0.0 and O.log2f(2.0) == 1.0 and O.log2f(0.5) == -1.0
    assert O.exp2f(0.0) == 1.0 and O.exp2f(3.0) == 8.0 and O.exp2f(-126.0) == np.float32(2.0 ** -126)
    assert O.exp2f(-200.0) == 0.0 and O.exp2f(200.0) == np.inf and O.exp2f(-np.inf) == 0.0
    assert O.powf(0.0, 0.7) == 0.0            # phi on an edge -> window 0
    assert np.isnan(O.powf(-0.01, 0.7))       # outside the triangle -> NaN -> clamped to 0 by max(0, .)
    assert O.powf(1.0, 3.3) == 1.0


def test_sincos_against_libm():
    xs = np.linspace(0.0, 2 * math.pi, 4001, dtype=np.float32)
    err = 0.0
    for x in xs:
        s, c = O.sincosf(x)
        err = max(err, abs(float(s) - math.sin(float(x))), abs(float(c) - math.cos(float(x))))
    assert err <= 3e-7, err


# --------------------------------------------------------------- SH colour
SQPI = math.sqrt(math.pi)
C0 = 1 / (2 * SQPI)
C1 = math.sqrt(3) / (2 * SQPI)


def _sh_basis(f64, deg, d):
    """Basis values Y_k(d) as seen through the oracle: the colour of a one-hot
    coefficient set minus the +0.5 offset (single channel)."""
    M = (deg + 1) ** 2
    vals = []
    for k in range(M):
        sh = np.zeros((16, 3))
        sh[k, 0] = 1.0
        sh[k, 1] = -1.0
        rgb, _ = O.sh_color(sh, deg, d, f64=f64)
        vals.append(rgb[0] - 0.5 if rgb[0] > 0 else -(rgb[1] - 0.5))
    return np.array(vals, np.float64)


def test_sh_constants_closed_form():
    d = np.array([0.3, -0.5, 0.81])
    d = d / np.linalg.norm(d)
    x, y, z = d
    expected = [C0, -C1 * y, C1 * z, -C1 * x,
                0.5 * math.sqrt(15 / math.pi) * x * y, -0.5 * math.sqrt(15 / math.pi) * y * z,
                0.25 * math.sqrt(5 / math.pi) * (3 * z * z - 1), -0.5 * math.sqrt(15 / math.pi) * x * z,
                0.25 * math.sqrt(15 / math.pi) * (x * x - y * y),
                -0.25 * math.sqrt(35 / (2 * math.pi)) * y * (3 * x * x - y * y),
                0.5 * math.sqrt(105 / math.pi) * x * y * z,
                -0.25 * math.sqrt(21 / (2 * math.pi)) * y * (5 * z * z - 1),
                0.25 * math.sqrt(7 / math.pi) * z * (5 * z * z - 3),
                -0.25 * math.sqrt(21 / (2 * math.pi)) * x * (5 * z * z - 1),
                0.25 * math.sqrt(105 / math.pi) * z * (x * x - y * y),
                -0.25 * math.sqrt(35 / (2 * math.pi)) * x * (x * x - 3 * y * y)]
    got = _sh_basis(True, 3, d * 2.7)  # unnormalised direction: the oracle normalises
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-12)
    got32 = _sh_basis(False, 3, d)
    np.testing.assert_allclose(got32, expected, rtol=0, atol=3e-6)


def test_sh_basis_orthonormal():
    # Gauss-Legendre in z (8 nodes) x 16 uniform phi: exact for the degree-6 products.
    zs, wz = np.polynomial.legendre.leggauss(8)
    G = np.zeros((16, 16))
    for z, w in zip(zs, wz):
        r = math.sqrt(1 - z * z)
        for j in range(16):
            ph = 2 * math.pi * j / 16
            Y = _sh_basis(True, 3, np.array([r * math.cos(ph), r * math.sin(ph), z]))
            G += np.outer(Y, Y) * w * (2 * math.pi / 16)
    np.testing.assert_allclose(G, np.eye(16), atol=1e-12)


def test_sh_clamp_and_offset():
    sh = np.zeros((16, 3), np.float32)
    sh[0] = [-10.0, 0.0, 1.0]
    rgb, cl = O.sh_color(sh, 0, np.array([0, 0, 1.0]))
    assert rgb[0] == 0.0 and cl[0] and not cl[1] and not cl[2]
    assert abs(rgb[1] - 0.5) < 1e-7 and abs(rgb[2] - (C0 + 0.5)) < 1e-6


# ----------------------------------------------------------- window function
def _bary_closed_form(V, p):
    """phi_final by barycentrics: min_i lambda_i * perimeter / |edge opposite vertex i|."""
    A, B, C = V
    T = np.stack([B - A, C - A], 1)
    l1, l2 = np.linalg.lstsq(T, p - A, rcond=None)[0]
    lam = np.array([1 - l1 - l2, l1, l2])
    opp = np.array([np.linalg.norm(C - B), np.linalg.norm(C - A), np.linalg.norm(B - A)])
    return float(np.min(lam * opp.sum() / opp))


@pytest.mark.parametrize("f64", [True, False])
def test_window_closed_form(f64):
    rng = np.random.Generator(np.random.PCG64(1))
    for _ in range(200):
        V = rng.standard_normal((3, 3)) * 0.1 + rng.uniform(-1, 1, 3)
        w = rng.dirichlet([2, 2, 2])
        p = w @ V
        if not f64:  # the closed forms are evaluated on the FP32-rounded inputs
            V = V.astype(np.float32).astype(np.float64)
            p = p.astype(np.float32).astype(np.float64)
        sigma = float(np.exp(rng.normal(0, 0.5)))
        out = O.window(V, p, sigma, f64=f64)
        pf = _bary_closed_form(V, p)
        tol = 1e-10 if f64 else 2e-4
        assert abs(out["phi_final"] - pf) <= tol * max(1.0, pf), (out["phi_final"], pf)
        assert abs(out["Cx"] - pf ** sigma) <= (1e-9 if f64 else 5e-4)
        n = np.cross(V[1] - V[0], V[2] - V[0])
        np.testing.assert_allclose(out["normal"], n / np.linalg.norm(n), atol=1e-12 if f64 else 2e-6)
        # incenter distance = -(inradius) = -2*area/perimeter
        per = sum(np.linalg.norm(V[i] - V[(i + 1) % 3]) for i in range(3))
        assert abs(out["dist"] + np.linalg.norm(n) / per) <= (1e-12 if f64 else 1e-6)


def test_window_special_points():
    V = np.array([[0, 0, 0], [1, 0, 0], [0.3, 0.8, 0.0]])
    a, b, c = (np.linalg.norm(V[1] - V[2]), np.linalg.norm(V[0] - V[2]), np.linalg.norm(V[0] - V[1]))
    inc = (a * V[0] + b * V[1] + c * V[2]) / (a + b + c)
    for sigma in (0.3, 1.0, 4.0):
        assert abs(O.window(V, inc, sigma, f64=True)["Cx"] - 1.0) < 1e-12
        assert O.window(V, 0.5 * (V[0] + V[1]), sigma, f64=True)["Cx"] == 0.0   # on an edge
        assert O.window(V, V[2], sigma, f64=True)["Cx"] == 0.0                   # at a vertex
    assert O.window(V, np.array([2.0, 2.0, 0.0]), 1.5, f64=True)["Cx"] == 0.0    # outside
    p = 0.2 * V[0] + 0.5 * V[1] + 0.3 * V[2]
    assert abs(O.window(V, p, 1e-6, f64=True)["Cx"] - 1.0) < 1e-5                # sigma -> 0


# ------------------------------------------------------------- intersection
def _exact_hit(o, d, V):
    """Exact rational ray/plane/barycentric test: (hit, t)."""
    F = lambda v: [Fraction(float(x)) for x in v]
    o, d = F(o), F(d)
    A, B, C = (F(v) for v in V)
    sub = lambda u, v: [u[i] - v[i] for i in range(3)]
    cross = lambda u, v: [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    dot = lambda u, v: sum(u[i] * v[i] for i in range(3))
    e1, e2 = sub(B, A), sub(C, A)
    n = cross(e1, e2)
    den = dot(n, d)
    if den == 0:
        return False, None, 0
    t = dot(n, sub(A, o)) / den
    p = [o[i] + t * d[i] for i in range(3)]
    s = [dot(n, cross(sub(B, A), sub(p, A))), dot(n, cross(sub(C, B), sub(p, B))), dot(n, cross(sub(A, C), sub(p, C)))]
    inside = all(x >= 0 for x in s) or all(x <= 0 for x in s)
    margin = min(abs(x) for x in s) / (dot(n, n))
    return inside, t, float(margin)


def test_intersection_against_exact_arithmetic():
    rng = np.random.Generator(np.random.PCG64(2))
    checked = 0
    for _ in range(600):
        V = (rng.standard_normal((3, 3)) * 0.3).astype(np.float32)
        o = (rng.standard_normal(3) * 2).astype(np.float32)
        target = (rng.dirichlet([1, 1, 1]) * 1.4 - 0.2) @ V
        d = (target - o).astype(np.float32)
        d /= np.float32(np.linalg.norm(d))
        inside, t_ex, margin = _exact_hit(o, d, V)
        if margin < 1e-4:
            continue  # too close to an edge for an FP32 answer to be fixed
        t = O.intersect(o, d, *V)
        assert (t is not None) == inside
        if inside:
            assert abs(float(t) - float(t_ex)) <= 1e-5 * max(1.0, abs(float(t_ex)))
        checked += 1
    assert checked > 300


def test_intersection_watertight_on_shared_edges():
    # a fan of triangles around a shared vertex; rays through the shared vertex and
    # through points of shared edges must hit at least one triangle
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(50):
        c = rng.uniform(-1, 1, 3).astype(np.float32)
        ring = [c + np.float32(0.3) * np.array([math.cos(a), math.sin(a), 0.05 * math.sin(3 * a)], np.float32)
                for a in np.linspace(0, 2 * math.pi, 7)[:-1]]
        tris = [(c, ring[i], ring[(i + 1) % 6]) for i in range(6)]
        o = (c + np.array([0.05, -0.02, 2.0], np.float32)).astype(np.float32)
        for tgt in [c] + [np.float32(0.5) * (c + r) for r in ring]:
            d = (tgt - o).astype(np.float32)
            hits = [O.intersect(o, d, *tr) for tr in tris]
            assert any(h is not None for h in hits)


def test_intersection_at_the_origin_is_plus_zero():
    # origin on the triangle's plane, inside it: t = 0 / det, whose sign follows the winding;
    # both windings report +0.0 (hit keys order t as unsigned bits, -0.0 would sort last)
    V = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    for tri in (V, V[::-1].copy()):
        for d in ([0, 0, 1], [0.1, 0.2, -1]):
            t = O.intersect([0.0, 0.0, 0.0], d, *tri)
            assert t is not None and float(t) == 0.0 and not np.signbit(np.float32(t))


def test_intersection_double_sided_and_parallel():
    V = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    t1 = O.intersect([0.2, 0.2, 0], [0, 0, 1], *V)
    t2 = O.intersect([0.2, 0.2, 2], [0, 0, -1], *V)
    assert t1 == 1.0 and t2 == 1.0
    assert O.intersect([0.2, 0.2, 0], [1, 0, 0], *V) is None       # parallel
    assert O.intersect([2.0, 2.0, 0], [0, 0, 1], *V) is None       # outside
