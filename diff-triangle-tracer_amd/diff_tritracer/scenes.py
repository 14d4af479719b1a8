"""Seeded synthetic triangle-splat scenes and pinhole cameras.

The benchmark workloads are synthetic (the reference ships no scenes, tests or
benchmarks). Arrays come out in the layout the tracer's drop-in API takes
(diff_tritracer/__init__.py of the reference, :54-139):

  triangles_points            float32 [3P, 3]  three consecutive vertices per triangle
  num_points_per_triangle     int32   [P]      all 3
  cumsum_of_points_per_triangle int32 [P]      3*i
  shs                         float32 [P, 16, 3]
  opacities, sigma            float32 [P, 1]   (opacity already activated)
  ray_o, ray_d                float32 [H, W, 3] (unit directions)

Generators use numpy's PCG64 with an explicit seed, so inputs are identical on
every host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class SplatScene:
    verts: np.ndarray      # [3P,3] f32
    npts: np.ndarray       # [P] i32
    cumsum: np.ndarray     # [P] i32
    shs: np.ndarray        # [P,16,3] f32
    opac: np.ndarray       # [P,1] f32
    sigma: np.ndarray      # [P,1] f32
    sh_degree: int = 3

    @property
    def P(self) -> int:
        return int(self.cumsum.shape[0])


def _attrs(rng, P, scene_verts):
    opac = (1.0 / (1.0 + np.exp(-rng.standard_normal(P)))).astype(np.float32)
    sigma = np.exp(0.5 * rng.standard_normal(P)).astype(np.float32)
    shs = np.empty((P, 16, 3), np.float32)
    shs[:, 0, :] = 0.5 * rng.standard_normal((P, 3))
    shs[:, 1:, :] = 0.05 * rng.standard_normal((P, 15, 3))
    return SplatScene(
        verts=np.ascontiguousarray(scene_verts.reshape(-1, 3), np.float32),
        npts=np.full(P, 3, np.int32),
        cumsum=(3 * np.arange(P, dtype=np.int64)).astype(np.int32),
        shs=shs,
        opac=opac.reshape(P, 1),
        sigma=sigma.reshape(P, 1),
    )


def uniform_scene(P: int = 10_000, seed: int = 0, offset_std: float = 0.03) -> SplatScene:
    """Small config: centres U[-1,1]^3, vertices = centre + N(0, offset_std) offsets,
    opacity = sigmoid(N(0,1)), sigma = exp(N(0,0.5)), SH DC ~ N(0,0.5), rest ~ N(0,0.05)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-1.0, 1.0, (P, 1, 3))
    verts = centres + offset_std * rng.standard_normal((P, 3, 3))
    return _attrs(rng, P, verts.astype(np.float32))


def _random_rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def clustered_scene(P: int = 1_000_000, clusters: int = 64, seed: int = 1, edge_mu: float = -4.0,
                    edge_s: float = 0.5, extent: float = 2.0) -> SplatScene:
    """Scale config: centres from a mixture of anisotropic Gaussian clusters inside
    [-extent, extent]^3; triangle edge lengths lognormal(edge_mu, edge_s); random
    orientation; opacity, sigma and SH as in the small config."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cc = rng.uniform(-0.75 * extent, 0.75 * extent, (clusters, 3))
    axes = np.exp(rng.uniform(math.log(0.03), math.log(0.4), (clusters, 3))) * extent / 2.0
    rot = _random_rotations(rng, clusters)
    weights = rng.uniform(0.5, 1.5, clusters)
    weights /= weights.sum()
    lab = rng.choice(clusters, size=P, p=weights)
    z = rng.standard_normal((P, 3)) * axes[lab]
    centres = cc[lab] + np.einsum("pij,pj->pi", rot[lab], z)
    centres = np.clip(centres, -extent, extent)
    # equilateral-ish triangle of edge ~ lognormal, random plane, jittered angles
    edge = np.exp(edge_mu + edge_s * rng.standard_normal(P))
    circ = edge / math.sqrt(3.0)
    frame = _random_rotations(rng, P)
    u, v = frame[:, :, 0], frame[:, :, 1]
    base = rng.uniform(0, 2 * math.pi, P)
    verts = np.empty((P, 3, 3), np.float64)
    for j in range(3):
        th = base + 2 * math.pi * j / 3 + 0.25 * rng.standard_normal(P)
        verts[:, j, :] = centres + circ[:, None] * (np.cos(th)[:, None] * u + np.sin(th)[:, None] * v)
    return _attrs(rng, P, verts.astype(np.float32))


def camera_basis(width: int, height: int, eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg: float = 60.0):
    """Look-at pinhole basis in double: eye, forward, right, up (unit), tan of the half
    field of view in x (fov_deg is horizontal) and y."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    if np.linalg.norm(right) < 1e-8:
        right = np.cross(fwd, np.array([1.0, 0.0, 0.0]))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    tx = math.tan(math.radians(fov_deg) / 2)
    ty = tx * height / width
    return dict(eye=eye, forward=fwd, right=right, up=upv, tan_x=tx, tan_y=ty)


def look_at_rays(width: int, height: int, eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                 fov_deg: float = 60.0):
    """Pinhole camera: ray_o, ray_d as float32 [H, W, 3]; unit directions through pixel
    centres; fov_deg is the horizontal field of view."""
    b = camera_basis(width, height, eye, target, up, fov_deg)
    fwd, right, upv = b["forward"], b["right"], b["up"]
    xs = (2.0 * (np.arange(width) + 0.5) / width - 1.0) * b["tan_x"]
    ys = (1.0 - 2.0 * (np.arange(height) + 0.5) / height) * b["tan_y"]
    d = fwd[None, None, :] + xs[None, :, None] * right[None, None, :] + ys[:, None, None] * upv[None, None, :]
    d /= np.sqrt((d[..., 0:1] * d[..., 0:1] + d[..., 1:2] * d[..., 1:2]) + d[..., 2:3] * d[..., 2:3])
    o = np.broadcast_to(b["eye"], d.shape)
    return np.ascontiguousarray(o, np.float32), np.ascontiguousarray(d, np.float32)


def orbit_eye(radius: float, seed: int, index: int = 0):
    """A deterministic camera position on a sphere of the given radius."""
    rng = np.random.Generator(np.random.PCG64(seed * 1000 + index))
    v = rng.standard_normal(3)
    v /= np.linalg.norm(v)
    return radius * v


def small_config(seed: int = 0, width: int = 256, height: int = 256):
    """10k splats, 256x256 camera at radius 3 looking at the origin, fov 60 deg."""
    scene = uniform_scene(10_000, seed)
    eye = 3.0 * np.array([0.3, 0.4, 1.0]) / np.linalg.norm([0.3, 0.4, 1.0])
    ro, rd = look_at_rays(width, height, eye)
    return scene, ro, rd


def scale_config(seed: int = 1, P: int = 1_000_000, width: int = 1920, height: int = 1080, clusters: int = 64):
    """1M clustered splats, one 1920x1080 camera at radius 4 looking at the origin."""
    scene = clustered_scene(P, clusters, seed)
    eye = 4.0 * np.array([0.35, 0.3, 1.0]) / np.linalg.norm([0.35, 0.3, 1.0])
    ro, rd = look_at_rays(width, height, eye)
    return scene, ro, rd


MULTIVIEW_SIZES = [(640, 480), (800, 600), (1024, 768), (1280, 720), (1280, 720), (1600, 900), (1920, 1080),
                   (1920, 1080)]


def multiview_eyes(n: int, seed: int = 2, rmin: float = 3.0, rmax: float = 5.0):
    """Positions of n cameras on random poses on a sphere of radius [rmin, rmax]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [orbit_eye(rng.uniform(rmin, rmax), seed, i) for i in range(n)]


def multiview_bases(sizes, seed: int = 2, rmin: float = 3.0, rmax: float = 5.0, fov_deg: float = 60.0):
    """camera_basis() of every camera of multiview_rays (same poses)."""
    eyes = multiview_eyes(len(sizes), seed, rmin, rmax)
    return [camera_basis(w, h, e, fov_deg=fov_deg) for (w, h), e in zip(sizes, eyes)]


def multiview_rays(sizes, seed: int = 2, rmin: float = 3.0, rmax: float = 5.0, fov_deg: float = 60.0):
    """Cameras of uneven resolution on random poses on a sphere of radius [rmin, rmax],
    looking at the origin. Returns (widths, heights, ray_o [sum R, 3], ray_d [sum R, 3])
    with the cameras' rays concatenated in order."""
    os_, ds_ = [], []
    for (w, h), eye in zip(sizes, multiview_eyes(len(sizes), seed, rmin, rmax)):
        ro, rd = look_at_rays(w, h, eye, fov_deg=fov_deg)
        os_.append(ro.reshape(-1, 3))
        ds_.append(rd.reshape(-1, 3))
    return ([w for w, _ in sizes], [h for _, h in sizes], np.ascontiguousarray(np.concatenate(os_)),
            np.ascontiguousarray(np.concatenate(ds_)))


def multiview_config(seed: int = 2, P: int = 2_000_000, sizes=None):
    """Batched multi-view step: 2M clustered splats, 8 cameras of uneven resolution
    (640x480 ... 1920x1080, 9,004,032 rays) on random poses at radius 3-5."""
    scene = clustered_scene(P, 64, seed)
    w, h, ro, rd = multiview_rays(sizes or MULTIVIEW_SIZES, seed)
    return scene, w, h, ro, rd
