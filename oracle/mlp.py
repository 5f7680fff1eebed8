"""Float64 numpy restatement of the learned-sort network (test
infrastructure only): FP32 MLP 1 -> 64 -> 64 -> 1 with ReLU, mean squared
error, parameter layout W1[64] b1[64] W2[64][64] b2[64] W3[64] b3.  No
counterpart in the reference; its analytic gradient is pinned by central
finite differences in tests/test_oracle_mlp.py."""
import numpy as np

H = 64
NPARAM = H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w1, b1 = p[0:64], p[64:128]
    w2 = p[128:128 + 4096].reshape(64, 64)  # [out j][in i]
    b2 = p[4224:4288]
    w3 = p[4288:4352]
    b3 = p[4352]
    return w1, b1, w2, b2, w3, b3


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    z1 = np.outer(x, w1) + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3


def loss(p, x, t):
    e = forward(p, x) - np.asarray(t, np.float64)
    return float(np.mean(e * e))


def grad(p, x, t):
    """Analytic gradient of loss(p, x, t) with respect to p."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    S = len(x)
    z1 = np.outer(x, w1) + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    dy = 2.0 * (y - t) / S
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    gw2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ w2) * (z1 > 0)
    gw1 = d1.T @ x
    gb1 = d1.sum(0)
    return np.concatenate([gw1, gb1, gw2.ravel(), gb2, gw3, [gb3]])
