"""Pins of the FP64 proposal-network oracle (CPU only): gradient vs central finite
differences, encoding vs closed forms, bf16 rounding vs the bit patterns of the
GPU's float->bf16 conversion, Adam's first step vs its closed form, and the batch
sampler vs the published splitmix64 sequence."""
import math

import numpy as np

import proposal as PR


def test_bf16_rounding_matches_device_conversion():
    # bit patterns produced by the gfx950 float->bf16 conversion for these inputs
    x = np.array([1.0 + 1.0 / 256, 1.0 + 3.0 / 256, 1.0 + 1.0 / 512, -2.5, 1e-40, 3.0e38], np.float32)
    assert list(PR.bf16_bits(x)) == [0x3F80, 0x3F82, 0x3F80, 0xC020, 0x0001, 0x7F62]


def test_encoding_closed_form():
    o = np.array([[2.0, 0.0, -4.0]])
    d = np.array([[0.0, 0.6, 0.8]])
    x = PR.encode(o, d, enc_scale=8.0)
    assert x.shape == (1, 102)
    np.testing.assert_allclose(x[0, :6], [0.25, 0.0, -0.5, 0.0, 0.6, 0.8])

    def feat(dim, f, cos):
        return x[0, 6 + 16 * dim + 2 * f + cos]

    assert abs(feat(0, 0, 0) - math.sqrt(0.5)) < 1e-15 and abs(feat(0, 0, 1) - math.sqrt(0.5)) < 1e-15
    assert abs(feat(0, 1, 0) - 1.0) < 1e-15 and abs(feat(0, 1, 1)) < 1e-15
    assert abs(feat(2, 0, 0) + 1.0) < 1e-15
    assert abs(feat(5, 3, 0) - math.sin(8 * math.pi * 0.8)) < 1e-15
    assert np.all(x[0, 6 + 16 * 1:6 + 16 * 2:2] == 0.0)  # sin of 0


def _problem(seed, B=96):
    rng = np.random.Generator(np.random.PCG64(seed))
    params = PR.pack({name: rng.normal(0, 1.0 / math.sqrt(shp[-1]), shp) if name[0] == "W" else
                      rng.normal(0, 0.1, shp) for name, shp in PR.SHAPES})
    o = rng.normal(0, 3.0, (B, 3))
    d = rng.normal(0, 1.0, (B, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    iv = np.empty((B, 4))
    iv[:, 0] = rng.uniform(1.0, 6.0, B)
    iv[:, 1] = iv[:, 0] + rng.uniform(0.0, 1.0, B)
    iv[:, 2] = np.where(rng.uniform(size=B) < 0.5, iv[:, 1] + 0.1, np.inf)
    nohit = rng.uniform(size=B) < 0.2
    iv[nohit, :3] = np.inf
    return params, o, d, iv


def test_gradient_matches_finite_differences():
    params, o, d, iv = _problem(0)
    loss, g = PR.loss_grad(params, o, d, iv, use_bf16=False)
    assert g.shape == (PR.N_PARAMS,) and loss > 0
    rng = np.random.Generator(np.random.PCG64(1))
    idx = np.concatenate([rng.choice(PR.N_PARAMS, 300, replace=False),
                          np.arange(PR.N_PARAMS - 130, PR.N_PARAMS)])  # all of W4, b4 and some of b3
    h = 1e-6
    ok = 0
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (PR.loss_grad(pp, o, d, iv, use_bf16=False)[0] - PR.loss_grad(pm, o, d, iv, use_bf16=False)[0]) / (2 * h)
        ok += abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd)
    # a central difference that straddles a ReLU or pinball kink may miss; allow 1%
    assert ok >= 0.99 * len(idx), (ok, len(idx))


def test_bf16_emulation_stays_close_to_fp64():
    params, o, d, iv = _problem(2)
    l64, g64 = PR.loss_grad(params, o, d, iv, use_bf16=False)
    l16, g16 = PR.loss_grad(params, o, d, iv, use_bf16=True)
    assert abs(l16 - l64) < 2e-2 * abs(l64)
    cos = g16 @ g64 / np.linalg.norm(g16) / np.linalg.norm(g64)
    assert cos > 0.99


def test_adam_first_step_closed_form():
    rng = np.random.Generator(np.random.PCG64(3))
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    z = np.zeros(50)
    p1, m1, v1 = PR.adam(p, z, z, g, 1)
    hp = PR.DEFAULT_HP
    np.testing.assert_allclose(p1, p - hp["lr"] * g / (np.abs(g) + hp["eps"]), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m1, (1 - hp["beta1"]) * g)


def test_sampler_is_splitmix64():
    # splitmix64 from state 0: published first outputs
    assert PR.mix64(0) == 0xE220A8397B1DCDAF
    assert PR.mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    idx = PR.sampled_rays(seed=5, step=1, batch=64, n_rays=1000)
    assert idx.min() >= 0 and idx.max() < 1000 and len(set(idx.tolist())) > 50
