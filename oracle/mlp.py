"""Float64 oracle of the scheduler MLP -- TEST INFRASTRUCTURE.

Same parameter layout as the HIP kernels: W1[H][IN], b1[H], W2[H][H], b2[H],
W3[OUT][H], b3[OUT]; ReLU hidden layers; masked mean squared error
L = sum M (y - Y)^2 / sum M.  Gradients are derived by hand below and pinned
by central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 8, 64, 8
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)),
         ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (X, a1, h1, a2, h2)


def loss(p, X, Y, M):
    y, _ = forward(p, X)
    M = np.asarray(M, dtype=np.float64)
    return float(np.sum(M * (y - Y) ** 2) / np.sum(M))


def loss_grad(p, X, Y, M):
    w = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    M = np.asarray(M, dtype=np.float64)
    cnt = np.sum(M)
    L = float(np.sum(M * (y - Y) ** 2) / cnt)
    dy = 2.0 * M * (y - Y) / cnt
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ X
    g["b1"] = dh1.sum(0)
    return L, pack(g)
