"""CPU oracle of the cost-model MLP (test infrastructure only).

Restates the network the device kernels implement: in -> h1 -> h2 -> 1,
ReLU, parameters packed W1[h1][in], b1, W2[h2][h1], b2, W3[h2], b3, loss =
mean squared error.  The analytic gradient here is pinned by central finite
differences in float64 (tests/test_mlp.py); the device gradient is compared
against it in float32.
"""
import numpy as np


def unpack(p, i, h1, h2):
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + h1 * i].reshape(h1, i); o += h1 * i
    b1 = p[o:o + h1]; o += h1
    W2 = p[o:o + h2 * h1].reshape(h2, h1); o += h2 * h1
    b2 = p[o:o + h2]; o += h2
    W3 = p[o:o + h2]; o += h2
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def num_params(i, h1, h2):
    return h1 * i + h1 + h2 * h1 + h2 + h2 + 1


def forward(p, x, i, h1, h2):
    W1, b1, W2, b2, W3, b3 = unpack(p, i, h1, h2)
    a1 = x @ W1.T + b1
    z1 = np.maximum(a1, 0)
    a2 = z1 @ W2.T + b2
    z2 = np.maximum(a2, 0)
    return z2 @ W3 + b3, (a1, z1, a2, z2)


def loss(p, x, t, i, h1, h2):
    y, _ = forward(p, x, i, h1, h2)
    return float(np.mean((y - t) ** 2))


def grad(p, x, t, i, h1, h2):
    W1, b1, W2, b2, W3, b3 = unpack(p, i, h1, h2)
    y, (a1, z1, a2, z2) = forward(p, x, i, h1, h2)
    n = x.shape[0]
    dy = 2.0 * (y - t) / n
    gW3 = z2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, W3) * (a2 > 0)
    gW2 = d2.T @ z1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])


def init(i, h1, h2, seed=0, dtype=np.float32):
    """He-uniform initialisation."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((i, h1), (h1, h2), (h2, 1)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out) + (0.01 if fan_out > 1 else 0.0))
    return np.concatenate(parts).astype(dtype)
