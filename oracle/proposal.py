"""FP64 restatement of the learned depth-interval proposal -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's reporting may import this module; the product
(libtritracer_hip.so) never calls it. It restates, in numpy float64, the network
the HIP kernels in csrc/tt_proposal.hip run (there is no counterpart in the
reference: the OptiX tracer has no proposal stage):

  x   = enc(o, d): (o / enc_scale, d), then sin and cos of 2^f pi v, f = 0..7,
        for each of those six values v                                  [102]
  h1  = relu(W1 x + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)  [64 each]
  y   = W4 h3 + b4 = (start, end) / t_scale                             [2]
  loss per ray = pinball(y0 - u_start; tau_start) + pinball(y1 - u_end; tau_end),
        mean over the batch; targets u from the tracer's emitted interval:
        rays with a blended hit: (t_first, t_term or else t_last) / t_scale,
        others: (t_skip, t_skip) / t_scale
  Adam (bias-corrected) on the fp32 master parameters.

With bf16=True the values the GPU rounds to bfloat16 (round to nearest even) are
rounded here at the same points: the encoded input, the weights fed to the matrix
cores, the hidden activations and the backpropagated hidden gradients used as
matrix-core operands. Everything else is float64.

Parity pins (tests/test_oracle_proposal.py): the gradient against central finite
differences of the loss (bf16 off), the encoding against closed-form values, the
bf16 rounding against the bit patterns the GPU's conversion produced for the same
inputs, and Adam against its closed form at step 1.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 102, 64, 2
SHAPES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_OUT, N_HID)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)

DEFAULT_HP = dict(enc_scale=8.0, t_scale=8.0, t_skip=16.0, tau_start=0.1, tau_end=0.9, lr=1e-3, beta1=0.9,
                  beta2=0.999, eps=1e-8)


def unpack(flat):
    flat = np.asarray(flat, np.float64).reshape(-1)
    assert flat.size == N_PARAMS
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = flat[off:off + n].reshape(shp)
        off += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), via float32; returns float64."""
    x32 = np.asarray(x, np.float64).astype(np.float32)
    b = x32.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    """bfloat16 bit patterns (uint16) of float32 inputs, round to nearest even."""
    return (bf16(x).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def encode(ray_o, ray_d, enc_scale=8.0):
    o = np.asarray(ray_o, np.float64).reshape(-1, 3)
    d = np.asarray(ray_d, np.float64).reshape(-1, 3)
    v = np.concatenate([o / enc_scale, d], 1)  # [R, 6]
    feats = [v]
    per_dim = []
    for j in range(6):
        for f in range(8):
            a = (2.0 ** f) * np.pi * v[:, j]
            per_dim += [np.sin(a), np.cos(a)]
    feats.append(np.stack(per_dim, 1))
    return np.concatenate(feats, 1)  # [R, 102]


def targets(interval, hp):
    """[R][4] interval_out of the tracer -> (u_start, u_end) per ray."""
    iv = np.asarray(interval, np.float64).reshape(-1, 4)
    has = np.isfinite(iv[:, 0])
    stop = np.where(np.isfinite(iv[:, 2]), iv[:, 2], iv[:, 1])
    us = np.where(has, iv[:, 0], hp["t_skip"]) / hp["t_scale"]
    ue = np.where(has, stop, hp["t_skip"]) / hp["t_scale"]
    return np.stack([us, ue], 1)


def forward(params, x, use_bf16=True):
    p = unpack(params)
    r = bf16 if use_bf16 else (lambda a: a)
    a0 = r(x)
    z1 = a0 @ r(p["W1"]).T + p["b1"]
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(p["W2"]).T + p["b2"]
    h2 = r(np.maximum(z2, 0.0))
    z3 = h2 @ r(p["W3"]).T + p["b3"]
    h3 = r(np.maximum(z3, 0.0))
    y = h3 @ r(p["W4"]).T + p["b4"]
    return y, (a0, h1, h2, h3)


def pinball(res, tau):
    return np.where(res > 0, (1.0 - tau) * res, -tau * res), np.where(res > 0, 1.0 - tau, np.where(res < 0, -tau, 0.0))


def loss_grad(params, ray_o, ray_d, interval, hp=None, use_bf16=True):
    """Mean pinball loss over the rays and its gradient w.r.t. the flat parameters."""
    hp = dict(DEFAULT_HP, **(hp or {}))
    x = encode(ray_o, ray_d, hp["enc_scale"])
    u = targets(interval, hp)
    B = x.shape[0]
    p = unpack(params)
    r = bf16 if use_bf16 else (lambda a: a)
    y, (a0, h1, h2, h3) = forward(params, x, use_bf16)
    l0, g0 = pinball(y[:, 0] - u[:, 0], hp["tau_start"])
    l1, g1 = pinball(y[:, 1] - u[:, 1], hp["tau_end"])
    loss = float((l0 + l1).sum() / B)
    dy = np.stack([g0, g1], 1) / B
    g = {}
    g["W4"] = dy.T @ h3
    g["b4"] = dy.sum(0)
    dh3 = (dy @ r(p["W4"])) * (h3 > 0)
    g["b3"] = dh3.sum(0)
    dh3r = r(dh3)
    g["W3"] = dh3r.T @ h2
    dh2 = (dh3r @ r(p["W3"])) * (h2 > 0)
    g["b2"] = dh2.sum(0)
    dh2r = r(dh2)
    g["W2"] = dh2r.T @ h1
    dh1 = (dh2r @ r(p["W2"])) * (h1 > 0)
    g["b1"] = dh1.sum(0)
    g["W1"] = r(dh1).T @ a0
    return loss, pack(g)


def adam(params, m, v, grad, t, hp=None):
    """One bias-corrected Adam step at step t (1-based); returns (params, m, v)."""
    hp = dict(DEFAULT_HP, **(hp or {}))
    m = hp["beta1"] * m + (1 - hp["beta1"]) * grad
    v = hp["beta2"] * v + (1 - hp["beta2"]) * grad * grad
    mh = m / (1 - hp["beta1"] ** t)
    vh = v / (1 - hp["beta2"] ** t)
    return params - hp["lr"] * mh / (np.sqrt(vh) + hp["eps"]), m, v


def mix64(z):
    """splitmix64 finaliser (the GPU's batch sampler), on Python ints."""
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def sampled_rays(seed, step, batch, n_rays):
    """Ray indices the GPU trainer draws at Adam step `step` (1-based)."""
    M = (1 << 64) - 1
    sd = (seed * 0xD1B54A32D192ED03 + step) & M
    return np.array([mix64(sd ^ mix64(i)) % n_rays for i in range(batch)], np.int64)
