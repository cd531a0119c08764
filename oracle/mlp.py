"""CPU restatement of the tuner's cost-model network (test infrastructure).

Only tests/ import this module, as the reference for the HIP network in
threadcoarsenedstencilgenerator_amd/csrc/tcsg_mlp.hip.  The network has no
counterpart in the reference (hip/benchmark.cpp:221-263 measures every block
shape; nothing there predicts one), so the arithmetic here is pinned by the
mathematics instead of reference vectors: central finite differences of the
loss for every gradient, closed forms for the first Adam step and for a
network whose hidden units are all active, and a hand-computed tiny case.

Float64 throughout; the HIP network computes in float32, so comparisons
carry a tolerance.

Parameter layout (shared with the HIP code): W1[h][in], b1[h], W2[h][h],
b2[h], W3[1][h], b3[1], flattened in that order.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def num_params(in_dim: int, hidden: int) -> int:
    return hidden * in_dim + hidden + hidden * hidden + hidden + hidden + 1


def split(p: np.ndarray, in_dim: int, hidden: int):
    """Views (W1, b1, W2, b2, W3, b3) into the flat parameter vector."""
    h, i = hidden, in_dim
    o = np.cumsum([0, h * i, h, h * h, h, h, 1])
    return (p[o[0]:o[1]].reshape(h, i), p[o[1]:o[2]], p[o[2]:o[3]].reshape(h, h), p[o[3]:o[4]],
            p[o[4]:o[5]].reshape(1, h), p[o[5]:o[6]])


def _splitmix(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def init_params(in_dim: int, hidden: int, seed: int) -> np.ndarray:
    """Uniform weights U[-1/sqrt(fan_in), 1/sqrt(fan_in)) from one
    SplitMix64 stream (W1, then W2, then W3), zero biases, rounded to float32
    as the HIP network stores them."""
    p = np.zeros(num_params(in_dim, hidden), dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = split(p, in_dim, hidden)
    s = seed & _M64
    for W, fan in ((W1, in_dim), (W2, hidden), (W3, hidden)):
        lim = np.float32(np.float32(1.0) / np.sqrt(np.float32(fan)))
        flat = W.reshape(-1)
        for k in range(flat.size):
            s, z = _splitmix(s)
            u = (z >> 11) * 2.0 ** -53 * 2.0 - 1.0
            flat[k] = float(np.float32(u * float(lim)))
    return p


def forward(p, x, in_dim, hidden):
    """Returns (prediction[B], cache) for inputs x[B][in]."""
    W1, b1, W2, b2, W3, b3 = split(np.asarray(p, dtype=np.float64), in_dim, hidden)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = (h2 @ W3.T + b3)[:, 0]
    return y, (x, h1, h2)


def loss(p, x, y, in_dim, hidden) -> float:
    pred, _ = forward(p, x, in_dim, hidden)
    return float(np.mean((pred - np.asarray(y, dtype=np.float64)) ** 2))


def gradients(p, x, y, in_dim, hidden):
    """(grad[num_params], loss) of the mean-squared error, by backpropagation."""
    p = np.asarray(p, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = split(p, in_dim, hidden)
    pred, (x, h1, h2) = forward(p, x, in_dim, hidden)
    B = x.shape[0]
    d = pred - np.asarray(y, dtype=np.float64)
    g = np.zeros_like(p)
    gW1, gb1, gW2, gb2, gW3, gb3 = split(g, in_dim, hidden)
    dout = (2.0 / B) * d                      # dL/dpred
    gW3[...] = dout[None, :] @ h2
    gb3[...] = dout.sum()
    d2 = dout[:, None] * W3                   # dL/dh2
    d2 = d2 * (h2 > 0)
    gW2[...] = d2.T @ h1
    gb2[...] = d2.sum(axis=0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1[...] = d1.T @ x
    gb1[...] = d1.sum(axis=0)
    return g, float(np.mean(d * d))


class Adam:
    """Bias-corrected Adam with the HIP network's constants."""

    def __init__(self, n, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)


def finite_difference(p, x, y, in_dim, hidden, idx, h=1e-6):
    """Central difference of the loss along parameter indices idx."""
    p = np.asarray(p, dtype=np.float64)
    out = np.empty(len(idx))
    for k, j in enumerate(idx):
        a = p.copy()
        a[j] += h
        b = p.copy()
        b[j] -= h
        out[k] = (loss(a, x, y, in_dim, hidden) - loss(b, x, y, in_dim, hidden)) / (2 * h)
    return out
