"""The oracle's envlight hemisphere sampler (envlight.cu:25-51: an LCG draws (u, v), the
direction is cosine-weighted around the sphere normal) pinned by what fixes it
independently of its own code: the published sequence of the LCG it uses (Numerical
Recipes' quick generator, x' = 1664525 x + 1013904223 mod 2^32, from x = 0: 1013904223,
1196435762, 3519870697, 2868466484) and the moments of the cosine-weighted hemisphere
distribution (E[cos] = 2/3, E[cos^2] = 1/2, no mean tangential component)."""
import numpy as np
import pytest

import oracle as O

# x_1 .. x_4 of the generator from state 0 (published sequence)
LCG_FROM_ZERO = [1013904223, 1196435762, 3519870697, 2868466484]


def test_lcg_sequence_sets_first_polar_angle():
    # around +z the local frame is the world frame up to a rotation about z, so the first
    # sample's cos(theta) is sqrt(u1) with u1 = the low 24 bits of x_1 / 2^24, and the
    # second sample's is sqrt(u3) (each sample draws two numbers)
    w = O.cos_hemi(0, [0.0, 0.0, 1.0], 2)
    for k, x in ((0, LCG_FROM_ZERO[0]), (1, LCG_FROM_ZERO[2])):
        u = (x & 0xFFFFFF) / 16777216.0
        assert abs(float(w[k, 2]) ** 2 - u) < 2e-6
    # azimuth of the first sample: phi = 2 pi v1 with v1 from x_2, measured in the
    # tangent frame (t, b) = (normalize(up x n), n x t) with up = +y for n = +z
    v = (LCG_FROM_ZERO[1] & 0xFFFFFF) / 16777216.0
    t, b = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    phi = np.arctan2(w[0] @ b, w[0] @ t) % (2 * np.pi)
    assert abs(phi - 2 * np.pi * v) < 1e-4


@pytest.mark.parametrize("normal", [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0], [0.3, -0.5, 0.8],
                                    [0.02, 0.01, 0.9997]])
def test_cosine_weighted_moments(normal):
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    w = O.cos_hemi(2024, n.astype(np.float32), 200_000).astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(w, axis=1) - 1.0) < 1e-5)
    c = w @ n
    assert c.min() >= -1e-6  # every direction in the hemisphere around the normal
    # moments of p(w) = cos(theta) / pi (standard errors ~ 5e-4)
    assert abs(c.mean() - 2.0 / 3.0) < 4e-3
    assert abs((c * c).mean() - 0.5) < 4e-3
    tang = w - c[:, None] * n[None, :]
    assert np.linalg.norm(tang.mean(axis=0)) < 6e-3


def test_uniform_hemisphere_would_fail():
    # the moment test separates cosine-weighted from uniform sampling (E[cos] = 1/2 there)
    rng = np.random.default_rng(0)
    z = rng.uniform(0.0, 1.0, 200_000)
    assert abs(z.mean() - 2.0 / 3.0) > 0.1
