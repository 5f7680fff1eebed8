"""Pins the float64 MLP restatement (oracle/mlp.py) before the GPU kernels
are checked against it: the analytic gradient must match central finite
differences of the loss, and a zero-weight network must output b3."""
import numpy as np

from oracle import mlp


def _params(rng):
    p = rng.standard_normal(mlp.NPARAM) * 0.3
    p[0:64] = np.abs(p[0:64]) + 0.1
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.random(40)
    t = np.sort(rng.random(40))
    g = mlp.grad(p, x, t)
    h = 1e-6
    idx = np.concatenate([rng.choice(64, 8, replace=False),            # W1
                          64 + rng.choice(64, 8, replace=False),       # b1
                          128 + rng.choice(4096, 24, replace=False),   # W2
                          4224 + rng.choice(64, 8, replace=False),     # b2
                          4288 + rng.choice(64, 8, replace=False),     # W3
                          [4352]])                                     # b3
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss(pp, x, t) - mlp.loss(pm, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_special_cases():
    p = np.zeros(mlp.NPARAM)
    p[4352] = 0.25
    assert np.allclose(mlp.forward(p, np.linspace(0, 1, 5)), 0.25)
    # one path: h1 = relu(x), h2_0 = relu(2*h1_0), y = 3*h2_0 -> 6x
    p = np.zeros(mlp.NPARAM)
    p[0] = 1.0
    p[128 + 0 * 64 + 0] = 2.0
    p[4288] = 3.0
    assert np.allclose(mlp.forward(p, [0.5, -1.0]), [3.0, 0.0])
