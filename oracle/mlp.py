"""CPU restatement of the level predictor's network (TEST INFRASTRUCTURE ONLY): forward pass and
the gradient of the mean squared error, in float64 numpy, over the same parameter layout as the
HIP kernels (csrc/mlp.hip): [W1 (12x128) b1 | W2 (128x128) b2 | W3 (128x128) b3 | W4 (128x3) b4],
W input-major, SiLU hidden activations, inputs standardised as (x - mean) * scale.

It has no counterpart in libgrpp; it is pinned by central finite differences of its own loss
(tests/test_oracle_mlp.py)."""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 12, 128, 3
SHAPES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p: np.ndarray):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, mean, scale):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = (np.asarray(X, np.float64) - mean) * scale
    z1 = x @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    z3 = h2 @ W3 + b3
    h3 = silu(z3)
    y = h3 @ W4 + b4
    return y, (x, z1, h1, z2, h2, z3, h3)


def loss(p, X, T, mean, scale):
    y, _ = forward(p, X, mean, scale)
    return float(np.mean((y - T) ** 2))


def gradient(p, X, T, mean, scale):
    """d loss / d p for loss = mean over samples and outputs of (y - t)^2"""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (x, z1, h1, z2, h2, z3, h3) = forward(p, X, mean, scale)
    n = X.shape[0]
    dy = 2.0 * (y - np.asarray(T, np.float64)) / (n * NOUT)
    gW4, gb4 = h3.T @ dy, dy.sum(0)
    d3 = (dy @ W4.T) * silu_grad(z3)
    gW3, gb3 = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ W3.T) * silu_grad(z2)
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * silu_grad(z1)
    gW1, gb1 = x.T @ d1, d1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
