"""FP64 NumPy restatement of the warm-start U-Net (forward, backward, loss).

TEST INFRASTRUCTURE ONLY: the checker for the HIP BF16/MFMA kernels of
libtealeaf_hip.  The network has no counterpart in the reference; this file
fixes its definition:

  16-channel input (6 used: u_prev, density, energy, Kx, Ky and the residual
  of the previous-step guess, normalised)
  enc1: conv3x3 16->16, ReLU; conv3x3 16->16, ReLU            (skip1, HxW)
  avg-pool 2x2
  enc2: conv3x3 16->32, ReLU; conv3x3 32->32, ReLU            (skip2, H/2)
  avg-pool 2x2
  mid:  conv3x3 32->64, ReLU; conv3x3 64->64, ReLU            (H/4)
  nearest upsample x2, concat [up (64) | skip2 (32)] = 96
  dec2: conv3x3 96->32, ReLU; conv3x3 32->32, ReLU
  nearest upsample x2, concat [up (32) | skip1 (16)] = 48
  dec1: conv3x3 48->16, ReLU; conv3x3 16->16, ReLU
  out:  conv3x3 16->1 (linear)

Convolutions use zero padding.  Weights are W[tap, co, ci] with
tap = 3*(dy+1) + (dx+1), the padded layout the kernels use; only the first
`cin`/`cout` channels of the input and output layers are real.  The loss is
the mean over the batch's pixels of (out - target)^2.
"""
from __future__ import annotations

import numpy as np

# (cin_real, cin_pad, cout_real, cout_pad, relu)
LAYERS = [
    (6, 16, 16, 16, True), (16, 16, 16, 16, True),
    (16, 16, 32, 32, True), (32, 32, 32, 32, True),
    (32, 32, 64, 64, True), (64, 64, 64, 64, True),
    (96, 96, 32, 32, True), (32, 32, 32, 32, True),
    (48, 48, 16, 16, True), (16, 16, 16, 16, True),
    (16, 16, 1, 16, False),
]


def param_count() -> int:
    return sum(9 * ci * co + co for ci, _, co, _, _ in LAYERS)


def unflatten(flat: np.ndarray):
    """Flat real parameters -> list of (W[9, cout_pad, cin_pad], b[cout_pad])."""
    out, o = [], 0
    for ci, cip, co, cop, _ in LAYERS:
        w = np.zeros((9, cop, cip))
        n = 9 * co * ci
        w[:, :co, :ci] = flat[o:o + n].reshape(9, co, ci)
        o += n
        b = np.zeros(cop)
        b[:co] = flat[o:o + co]
        o += co
        out.append((w, b))
    assert o == len(flat)
    return out


def flatten_grads(grads) -> np.ndarray:
    parts = []
    for (ci, _, co, _, _), (dw, db) in zip(LAYERS, grads):
        parts.append(dw[:, :co, :ci].ravel())
        parts.append(db[:co])
    return np.concatenate(parts)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, as the library initialises them."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, _, co, _, relu in LAYERS:
        std = np.sqrt(2.0 / (9 * ci)) if relu else np.sqrt(1.0 / (9 * ci))
        parts.append(rng.standard_normal(9 * co * ci) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts)


def _im2col(x):
    B, H, W, C = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = np.empty((B, H, W, 9, C))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        cols[:, :, :, t, :] = xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]
    return cols


def conv(x, w, b):
    return np.einsum("bhwtc,toc->bhwo", _im2col(x), w, optimize=True) + b


def conv_back(x, w, dz):
    """Gradients of conv for dz = dL/d(conv output): (dx, dw, db)."""
    cols = _im2col(x)
    dw = np.einsum("bhwtc,bhwo->toc", cols, dz, optimize=True)
    db = dz.sum(axis=(0, 1, 2))
    # dx: transposed convolution = conv of dz with flipped, transposed weights
    wt = np.transpose(w[::-1], (0, 2, 1))  # [tap', ci, co] = w[8 - tap', co, ci]
    dx = np.einsum("bhwtc,toc->bhwo", _im2col(dz), wt, optimize=True)
    return dx, dw, db


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(params, x16):
    """x16: [B, H, W, 16] input (H, W multiples of 4).  Returns (out[B,H,W],
    cache)."""
    P = unflatten(params) if not isinstance(params, list) else params
    c = {}

    def layer(i, x):
        w, b = P[i]
        z = conv(x, w, b)
        y = np.maximum(z, 0.0) if LAYERS[i][4] else z
        c[i] = (x, z, y)
        return y

    e1 = layer(1, layer(0, x16))
    e2 = layer(3, layer(2, pool(e1)))
    m = layer(5, layer(4, pool(e2)))
    c2 = np.concatenate([up(m), e2], axis=-1)
    d2 = layer(7, layer(6, c2))
    c1 = np.concatenate([up(d2), e1], axis=-1)
    d1 = layer(9, layer(8, c1))
    out = layer(10, d1)
    return out[..., 0], (P, c)


def loss_and_grads(params, x16, target):
    """MSE loss and flat gradients w.r.t. the real parameters."""
    out, (P, c) = forward(params, x16)
    n = out.size
    diff = out - target
    loss = float(np.mean(diff * diff))
    grads = [None] * len(LAYERS)

    def back(i, dy):
        x, z, y = c[i]
        dz = dy * (z > 0) if LAYERS[i][4] else dy
        dx, dw, db = conv_back(x, P[i][0], dz)
        grads[i] = (dw, db)
        return dx

    dout = np.zeros(c[10][2].shape)
    dout[..., 0] = 2.0 * diff / n
    dd1 = back(8, back(9, back(10, dout)))
    dup1, de1_skip = dd1[..., :32], dd1[..., 32:]
    dd2 = back(6, back(7, up_back(dup1)))
    dup2, de2_skip = dd2[..., :64], dd2[..., 64:]
    dm = back(4, back(5, up_back(dup2)))
    de2 = de2_skip + pool_back(dm)
    dp1 = back(2, back(3, de2))
    de1 = de1_skip + pool_back(dp1)
    back(0, back(1, de1))
    return loss, flatten_grads(grads)
