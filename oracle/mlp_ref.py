"""NumPy restatement of the two-layer perceptron (oracle; test infrastructure
only): forward, mean-squared-error backward, Adam — the CPU checker of the HIP
kernels in ``csrc/smalgo_mlp.hip``.  Parameter vector order: W1 (nin×nh,
row-major), b1, W2 (nh×nout), b2."""
from __future__ import annotations

import numpy as np


def unpack(p, nin, nh, nout):
    o = 0
    W1 = p[o:o + nin * nh].reshape(nin, nh); o += nin * nh
    b1 = p[o:o + nh]; o += nh
    W2 = p[o:o + nh * nout].reshape(nh, nout); o += nh * nout
    b2 = p[o:o + nout]
    return W1, b1, W2, b2


def forward(p, X, nin, nh, nout):
    W1, b1, W2, b2 = unpack(p, nin, nh, nout)
    H = np.maximum(X @ W1 + b1, 0.0)
    return H @ W2 + b2, H


def loss(p, X, T, nin, nh, nout):
    Y, _ = forward(p, X, nin, nh, nout)
    return float(np.mean((Y - T) ** 2))


def gradient(p, X, T, nin, nh, nout):
    W1, b1, W2, b2 = unpack(p, nin, nh, nout)
    Y, H = forward(p, X, nin, nh, nout)
    dY = 2.0 * (Y - T) / Y.size
    gW2 = H.T @ dY
    gb2 = dY.sum(axis=0)
    dH = (dY @ W2.T) * (H > 0)
    gW1 = X.T @ dH
    gb1 = dH.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
