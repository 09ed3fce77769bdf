"""The CPU restatement of the level predictor's network (oracle/mlp.py) against what the
mathematics fixes: central finite differences of its own loss."""
import numpy as np

from oracle import mlp as M


def _case(seed, n=7):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.15, M.NPARAM)
    X = rng.normal(0.0, 2.0, (n, M.NIN))
    T = rng.normal(0.0, 1.0, (n, M.NOUT))
    mean, scale = rng.normal(0.0, 0.5, M.NIN), rng.uniform(0.5, 1.5, M.NIN)
    return p, X, T, mean, scale


def test_gradient_matches_central_differences():
    p, X, T, mean, scale = _case(0)
    g = M.gradient(p, X, T, mean, scale)
    rng = np.random.default_rng(1)
    # coordinates from every block: W1, b1, W2, b2, W3, b3, W4, b4
    bounds = np.cumsum([0] + [int(np.prod(s)) for s in M.SHAPES])
    idx = [int(rng.integers(bounds[k], bounds[k + 1])) for k in range(len(M.SHAPES)) for _ in range(4)]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss(p + e, X, T, mean, scale) - M.loss(p - e, X, T, mean, scale)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_silu_derivative():
    z = np.linspace(-8, 8, 33)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.allclose(fd, M.silu_grad(z), atol=1e-8)
