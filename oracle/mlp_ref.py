"""CPU reference of the surrogate network -- TEST INFRASTRUCTURE ONLY.

A float64 NumPy restatement of the network the library runs on the GPU
(144 -> 256 -> 256 -> 256 -> 12, SiLU; Gaussian negative log-likelihood on
standardised targets), used by tests/ to check the GPU forward pass and
gradients.  With emulate_bf16=True every operand that the GPU feeds to its
matrix cores is rounded to BF16 (round to nearest even) at the same place, so
GPU and reference differ only by accumulation order; with emulate_bf16=False
it is the exact network, which tests/test_mlp_reference.py pins against
central finite differences.  The network has no counterpart in the reference
repository.
"""
from __future__ import annotations

import numpy as np

SIZES = [(144, 256), (256, 256), (256, 256), (256, 12)]


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    """Flat parameters (W1, b1, ..., W4, b4; W row-major K x N) -> [(W, b)]."""
    out, o = [], 0
    for k, n in SIZES:
        w = np.asarray(p[o:o + k * n], np.float64).reshape(k, n)
        o += k * n
        b = np.asarray(p[o:o + n], np.float64)
        o += n
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x, x_mean, x_std, emulate_bf16=True):
    """Standardised outputs (n x 12) and the activations of each layer."""
    rnd = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    istd = (1.0 / np.asarray(x_std, np.float32)).astype(np.float32)
    xs = ((np.asarray(x, np.float32) - np.asarray(x_mean, np.float32)) * istd).astype(np.float32)
    h = rnd(xs)
    cache = {"in": [h], "z": []}
    for i, (w, b) in enumerate(layers):
        z = h @ (bf16(w) if emulate_bf16 else w) + b
        if i == len(layers) - 1:
            return z, cache
        cache["z"].append(z)
        h = rnd(_silu(z).astype(np.float32) if emulate_bf16 else _silu(z))
        cache["in"].append(h)


def loss_grad(p, x, t, x_mean, x_std, t_mean, t_std, emulate_bf16=True):
    """Mean Gaussian NLL 0.5 sum_k (s_k + (t'_k - mu_k)^2 exp(-s_k)) over the
    samples (mu = y[:, :6], s = y[:, 6:] clamped to +-20 in the value only) and
    its gradient in the flat parameter layout."""
    rnd = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    n = len(x)
    y, cache = forward(p, x, x_mean, x_std, emulate_bf16)
    tistd = (1.0 / np.asarray(t_std, np.float32)).astype(np.float32)
    tt = ((np.asarray(t, np.float32) - np.asarray(t_mean, np.float32)) * tistd).astype(np.float64)
    mu, s = y[:, :6], np.clip(y[:, 6:], -20.0, 20.0)
    e = np.exp(-s)
    d = tt - mu
    loss = float(np.sum(0.5 * (s + d * d * e)) / n)
    dy = np.concatenate([-d * e / n, 0.5 * (1.0 - d * d * e) / n], axis=1)
    layers = unpack(p)
    grads = [None] * len(layers)
    dz = dy
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        dzr = rnd(dz.astype(np.float32) if emulate_bf16 else dz)
        grads[i] = (cache["in"][i].T @ dzr, dz.sum(axis=0))
        if i == 0:
            break
        dh = dzr @ (bf16(w) if emulate_bf16 else w).T
        z = cache["z"][i - 1]
        sg = 1.0 / (1.0 + np.exp(-z))
        dz = dh * (sg * (1.0 + z * (1.0 - sg)))
    return loss, pack(grads)
