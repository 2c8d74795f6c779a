"""ORACLE — test infrastructure only.

FP64 numpy statement of the return-map warm-start network: 8 → 64 → 64 → 64 → 1, ReLU on
the hidden layers, linear output, mean-squared-error loss, Adam. Tests use it to check the
GPU training kernels' forward pass, gradients and optimiser step, and the bf16 MFMA
inference. It has no counterpart in the reference (the warm start is an addition); it is
pinned by a central finite-difference gradient check in tests/test_oracle_mlp.py.

Parameter order (the flat layout the GPU kernels share):
  W1 (64×8), b1 (64), W2 (64×64), b2 (64), W3 (64×64), b3 (64), W4 (1×64), b4 (1).
"""
from __future__ import annotations

import numpy as np

SIZES = [8, 64, 64, 64, 1]


def shapes():
    out = []
    for a, b in zip(SIZES[:-1], SIZES[1:]):
        out += [(b, a), (b,)]
    return out


def n_params():
    return sum(int(np.prod(s)) for s in shapes())


def unflatten(theta):
    parts, o = [], 0
    for s in shapes():
        k = int(np.prod(s))
        parts.append(theta[o:o + k].reshape(s))
        o += k
    return parts


def init(seed=0):
    rng = np.random.default_rng(seed)
    theta = []
    for (fan_out, fan_in), _ in zip(shapes()[0::2], shapes()[1::2]):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        theta.append(rng.uniform(-lim, lim, (fan_out, fan_in)).ravel())
        theta.append(np.zeros(fan_out))
    return np.concatenate(theta)


def forward(theta, x):
    """x (B, 8) normalised inputs -> (y (B,), cache)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(np.asarray(theta, np.float64))
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    h3 = np.maximum(h2 @ W3.T + b3, 0.0)
    y = (h3 @ W4.T + b4)[:, 0]
    return y, (x, h1, h2, h3)


def loss_and_grad(theta, x, t):
    """Mean squared error over the batch and its gradient (flat, same layout as theta)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(np.asarray(theta, np.float64))
    y, (x, h1, h2, h3) = forward(theta, x)
    B = x.shape[0]
    r = y - t
    loss = np.mean(r * r)
    dy = (2.0 / B) * r[:, None]
    gW4 = dy.T @ h3
    gb4 = dy.sum(0)
    dz3 = (dy @ W4) * (h3 > 0)
    gW3 = dz3.T @ h2
    gb3 = dz3.sum(0)
    dz2 = (dz3 @ W3) * (h2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def adam(theta, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected); returns new theta, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def features(strain_inc, eqps, temp, norm):
    """Raw (6 Voigt Δε, ε̄ᵖ, T) -> normalised network input with norm = (mean[8], std[8])."""
    x = np.concatenate([np.asarray(strain_inc), np.asarray(eqps)[:, None],
                        np.asarray(temp)[:, None]], 1)
    return (x - norm[0]) / norm[1]


def bf16(a):
    """Round to bfloat16 (round to nearest, ties to even) and back to float64: the top 16
    bits of the IEEE binary32 pattern after rounding the low 16."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def loss_and_grad_bf16(theta, x, t, rnd=bf16):
    """loss_and_grad with every matrix-product operand rounded by `rnd` where the GPU's bf16
    MFMA training step rounds it (sums in FP64 here, FP32 on the GPU): the inputs, hidden
    activations (after ReLU), the weight matrices W1-W4 of the forward pass and W2, W3 of
    the backward pass, the pre-activation gradients dz3, dz2, dz1, and dy where it multiplies
    h3ᵀ and sums into b4. Biases, the output, dy itself and dz3's factor W4 stay unrounded.
    With rnd = identity it is loss_and_grad."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unflatten(np.asarray(theta, np.float64))
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    xr = rnd(x)
    h1 = rnd(np.maximum(f32(xr @ rnd(W1).T + b1), 0.0))
    h2 = rnd(np.maximum(f32(h1 @ rnd(W2).T + b2), 0.0))
    h3 = rnd(np.maximum(f32(h2 @ rnd(W3).T + b3), 0.0))
    y = f32((h3 @ rnd(W4).T)[:, 0] + b4[0])
    B = x.shape[0]
    r = y - t
    loss = np.mean(r * r)
    dy = f32((2.0 / B) * r)[:, None]
    dyr = rnd(dy)
    gW4 = dyr.T @ h3
    gb4 = dyr.sum(0)
    dz3 = rnd(f32((dy @ W4) * (h3 > 0)))
    gW3 = dz3.T @ h2
    gb3 = dz3.sum(0)
    dz2 = rnd(f32((dz3 @ rnd(W3)) * (h2 > 0)))
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = rnd(f32((dz2 @ rnd(W2)) * (h1 > 0)))
    gW1 = dz1.T @ xr
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g
