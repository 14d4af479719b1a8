"""Closed-form pins for two oracle functions that GPU parity relies on (CPU only).

1. The refract program's direction rule (refract.cu:24-45, used by refract.cu:47-71)
   at oblique incidence. As written it takes sin^2(theta_i) = max(0, 1 - sqrt(cos_i))
   and sin^2(theta_t) = sin^2(theta_i) / sqrt(eta). For a ray entering the sphere
   cos_i < 0, sqrt(cos_i) is NaN and max(0, NaN) = 0, so the rule reduces to
       out = in / eta - (cos_i / eta + 1) n          (eta = 1.5f / 1.0003f)
   whatever the angle. For a ray leaving the sphere (camera inside) eta -> 1 / eta
   and n -> -n with cos_i kept positive; the fallback for sin^2(theta_t) >= 1 mirrors
   the ray about the normal facing it. Each case below is evaluated by hand in double
   and checked against the oracle's traced segment; each also asserts that physically
   correct Snell refraction gives a different answer, so replacing the quirk by correct
   Snell (or dropping the NaN clamp, or the flip) fails here.

2. The pinhole generator scenes.look_at_rays (the host reference of tt_camera_rays):
   centre ray = forward for odd sizes, corner rays at atan of the scaled half-field
   tangents, unit directions, row 0 at the top and column 0 on the left.
"""
import math

import numpy as np
import pytest

import oracle as O
from diff_tritracer import scenes

ETA = float(np.float32(1.5) / np.float32(1.0003))  # refract.cu:55-57, formed in float
STEP_EPS = 1e-9                                      # STEP_EPSILON, config.h:11


def _far_scene():
    """One invisible triangle far away: the trace only has to record its segment."""
    V = np.array([[100.0, 100.0, 100.0], [101.0, 100.0, 100.0], [100.0, 101.0, 100.0]], np.float32)
    return O.Scene(V, [3], [0], np.zeros((1, 16, 3), np.float32), 3, [0.5], [1.0])


def _traced_segment(o, d, sphere):
    S = _far_scene()
    log = O.Log(S, np.asarray([o], np.float32), np.asarray([d], np.float32),
                O.make_opts(cutoff=1e-4, variant="refract", extra=sphere))
    so, sd, role = log.segment_rays()
    assert so.shape[0] == 1
    return so[0].astype(np.float64), sd[0].astype(np.float64)


def _sphere_hit(o, d, r, c):
    """Hit distance of common.h:116-141 in double (the near root, else the far one)."""
    o, d, c = (np.asarray(v, np.float64) for v in (o, d, c))
    op = o - c
    a, b, cc = d @ d, op @ d, op @ op - r * r
    det = math.sqrt(b * b - cc * a)
    t = (-b - det) / a
    if t < STEP_EPS:
        t = (-b + det) / a
    return t


def _quirk_by_hand(inn, n, eta):
    cos_i = float(inn @ n)
    if cos_i > 0.0:
        eta, n = 1.0 / eta, -n
    s2i = 0.0 if cos_i < 0.0 else max(0.0, 1.0 - math.sqrt(cos_i))  # sqrt(<0) = NaN, max(0, NaN) = 0
    s2t = s2i / math.sqrt(eta)
    if s2t >= 1.0:
        return None
    cos_t = math.sqrt(max(0.0, 1.0 - s2t))
    return inn / eta + (-cos_i / eta - cos_t) * n


def _snell(inn, n, eta):
    """Physically correct refraction (pbrt's convention with the incident direction
    pointing towards the surface); None on total internal reflection."""
    cos_i = -float(inn @ n)
    if cos_i < 0.0:               # leaving: the ratio inverts, the normal faces the ray
        cos_i, n, eta = -cos_i, -n, 1.0 / eta
    s2t = (1.0 - cos_i * cos_i) / (eta * eta)
    if s2t >= 1.0:
        return None
    return inn / eta + (cos_i / eta - math.sqrt(1.0 - s2t)) * n


def _mirror(inn, n):
    rn = n if inn @ n < 0.0 else -n
    return inn - 2.0 * rn * (rn @ inn)


def _case(o, d, r=1.0, c=(0.0, 0.0, 0.0)):
    o, d, c = (np.asarray(v, np.float64) for v in (o, d, c))
    t = _sphere_hit(o, d, r, c)
    hp = o + t * d
    n = (hp - c) / np.linalg.norm(hp - c)
    inn = d / np.linalg.norm(d)
    return hp, n, inn


def test_refract_entering_40_degrees():
    th = math.radians(40.0)
    o, d = (math.sin(th), 0.0, 5.0), (0.0, 0.0, -1.0)
    hp, n, inn = _case(o, d)
    cos_i = float(inn @ n)
    assert abs(cos_i + math.cos(th)) < 1e-12          # incidence at 40 degrees
    want = inn / ETA - (cos_i / ETA + 1.0) * n         # the quirk's reduced form
    assert np.allclose(want, _quirk_by_hand(inn, n, ETA), atol=1e-15)
    go, gd = _traced_segment(o, d, [1.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(gd, want, atol=3e-6)
    np.testing.assert_allclose(go, hp + n * STEP_EPS, atol=3e-6)
    assert abs(np.linalg.norm(gd) - 1.0) > 1e-2        # not a unit vector: the quirk's signature
    snell = _snell(inn, n, ETA)
    assert snell is not None and abs(np.linalg.norm(snell) - 1.0) < 1e-12
    assert np.abs(gd - snell).max() > 5e-2             # correct Snell would fail this case


@pytest.mark.parametrize("deg", [15.0, 65.0])
def test_refract_entering_other_angles(deg):
    th = math.radians(deg)
    o, d = (0.2, 5.0, 0.3 + math.sin(th)), (0.0, -1.0, 0.0)
    c = (0.2, 0.0, 0.3)
    hp, n, inn = _case(o, d, 1.0, c)
    want = inn / ETA - (float(inn @ n) / ETA + 1.0) * n
    go, gd = _traced_segment(o, d, [1.0, *c])
    np.testing.assert_allclose(gd, want, atol=3e-6)
    assert np.abs(gd - _snell(inn, n, ETA)).max() > 1e-2


def test_refract_exiting_ray():
    # camera inside the sphere: the far root is taken, in . n > 0 (leaving), 30 degrees
    th = math.radians(30.0)
    d = np.array([math.sin(th), 0.0, -math.cos(th)])
    o = (0.0, 0.0, 0.0)
    hp, n, inn = _case(o, d)
    cos_i = float(inn @ n)
    assert cos_i > 0.99                                # leaving through the far side, near-normal
    want = _quirk_by_hand(inn, n, ETA)
    # by hand: eta' = 1 / eta, n' = -n, s2i = 1 - sqrt(cos_i), s2t = s2i / sqrt(eta')
    s2t = (1.0 - math.sqrt(cos_i)) / math.sqrt(1.0 / ETA)
    want2 = inn * ETA + (-cos_i * ETA - math.sqrt(1.0 - s2t)) * (-n)
    assert np.allclose(want, want2, atol=1e-15)
    go, gd = _traced_segment(o, d, [1.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(gd, want, atol=3e-6)
    np.testing.assert_allclose(go, hp + n * STEP_EPS, atol=3e-6)
    snell = _snell(inn, n, ETA)
    assert snell is not None and np.abs(gd - snell).max() > 1e-2


def test_refract_exiting_past_critical_angle_still_refracts():
    # 60 degrees from the normal on the way out: correct Snell has total internal
    # reflection (critical angle ~41.8 degrees), the quirk's s2t stays below 1
    th = math.radians(60.0)
    h = math.sin(th)                                   # offset of a ray parallel to x
    o, d = (0.0, h, 0.0), (1.0, 0.0, 0.0)
    hp, n, inn = _case(o, d)
    assert abs(float(inn @ n) - math.cos(th)) < 1e-12
    assert _snell(inn, n, ETA) is None
    want = _quirk_by_hand(inn, n, ETA)
    assert want is not None
    go, gd = _traced_segment(o, d, [1.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(gd, want, atol=3e-6)
    assert np.abs(gd - _mirror(inn, n)).max() > 1e-1  # what correct Snell (TIR) would give


def test_refract_grazing_exit_falls_back_to_mirror():
    # cos_i = 0.02 on the way out: s2i = 1 - sqrt(0.02) = 0.859 >= sqrt(1 / eta) = 0.817
    h = math.sqrt(1.0 - 0.02 ** 2)
    o, d = (0.0, h, 0.0), (1.0, 0.0, 0.0)
    hp, n, inn = _case(o, d)
    assert abs(float(inn @ n) - 0.02) < 1e-9
    assert _quirk_by_hand(inn, n, ETA) is None
    want = inn - 2.0 * n * float(inn @ n)              # reflect about -n (the normal facing the ray)
    go, gd = _traced_segment(o, d, [1.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(gd, want, atol=3e-6)
    np.testing.assert_allclose(go, hp + n * STEP_EPS, atol=3e-6)
    assert abs(np.linalg.norm(gd) - 1.0) < 1e-6


def test_refract_missed_sphere_keeps_the_ray():
    o, d = (3.0, 0.0, 5.0), (0.0, 0.0, -1.0)
    go, gd = _traced_segment(o, d, [1.0, 0.0, 0.0, 0.0])
    assert np.array_equal(go, np.asarray(o)) and np.array_equal(gd, np.asarray(d))


# --------------------------------------------------------------- pinhole camera
@pytest.mark.parametrize("W,H,fov,seed", [(101, 57, 60.0, 1), (33, 75, 45.0, 2), (1, 1, 90.0, 3)])
def test_look_at_centre_ray_is_forward(W, H, fov, seed):
    eye = scenes.orbit_eye(3.5, seed)
    ro, rd = scenes.look_at_rays(W, H, eye, fov_deg=fov)
    fwd = -eye / np.linalg.norm(eye)
    np.testing.assert_array_equal(ro[H // 2, W // 2], eye.astype(np.float32))
    np.testing.assert_allclose(rd[H // 2, W // 2], fwd, atol=1e-7)


@pytest.mark.parametrize("W,H,fov", [(1920, 1080, 60.0), (640, 480, 75.0), (97, 61, 45.0)])
def test_look_at_corner_angles(W, H, fov):
    # camera on +z looking at the origin with up = +y: right = +x, image row 0 at the top
    eye = np.array([0.0, 0.0, 3.0])
    ro, rd = scenes.look_at_rays(W, H, eye, fov_deg=fov)
    tx = math.tan(math.radians(fov) / 2.0)
    ty = tx * H / W
    ax = math.atan((1.0 - 1.0 / W) * tx)               # pixel centres sit half a pixel inside the edges
    ay = math.atan((1.0 - 1.0 / H) * ty)
    for (y, x), sx, sy in (((0, 0), -1, 1), ((0, W - 1), 1, 1), ((H - 1, 0), -1, -1), ((H - 1, W - 1), 1, -1)):
        d = rd[y, x].astype(np.float64)
        assert abs(math.atan2(d[0], -d[2]) - sx * ax) < 2e-7   # horizontal angle from the axis
        assert abs(math.atan2(d[1], -d[2]) - sy * ay) < 2e-7   # vertical angle
    # full horizontal spread between the outer pixel centres of the middle row
    a, b = rd[H // 2, 0].astype(np.float64), rd[H // 2, W - 1].astype(np.float64)
    spread = math.acos(min(1.0, a @ b / (np.linalg.norm(a) * np.linalg.norm(b))))
    want = 2.0 * math.atan((1.0 - 1.0 / W) * tx)
    if H % 2 == 1:
        assert abs(spread - want) < 1e-6


def test_look_at_unit_directions_and_shared_origin():
    eye = scenes.orbit_eye(4.0, 7)
    ro, rd = scenes.look_at_rays(160, 90, eye, fov_deg=60.0)
    n = np.linalg.norm(rd.astype(np.float64), axis=-1)
    assert np.abs(n - 1.0).max() < 2e-7
    assert np.array_equal(ro, np.broadcast_to(eye.astype(np.float32), ro.shape))
    # the directions of one row lie in one plane through the eye that contains `right`
    b = scenes.camera_basis(160, 90, eye)
    row = rd[17].astype(np.float64)
    nrm = np.cross(row[0], row[-1])
    assert np.abs(row @ nrm).max() < 1e-6 * np.linalg.norm(nrm) + 1e-7
    assert abs(b["right"] @ nrm) < 1e-6 * np.linalg.norm(nrm)
