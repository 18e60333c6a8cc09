"""CPU oracle of the learned match filter's network (test infrastructure only).

The same 264(288)-256-256-1 ReLU / sigmoid network and binary cross-entropy as
fieldstitch_amd/filter_mlp.py, in numpy float64, with the device's rounding points emulated:
BF16 weights (W1, W2) and activations (inputs, H1, H2, dZ2, dZ1) rounded to bfloat16 (round to
nearest even), FP32 master weights, biases and head weights.  `rounding=False` gives the plain
float64 network, whose backward pass is pinned by central finite differences.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    x = np.asarray(x, np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward(X, p, rounding=True):
    """X: (n, 288) inputs; p: dict W1 (288, 256), b1, W2 (256, 256), b2, w3 (256,), b3 (1,).
    Returns (prob, cache)."""
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))   # noqa: E731
    W1 = np.asarray(R(p["W1"]), np.float64)
    W2 = np.asarray(R(p["W2"]), np.float64)
    Xr = np.asarray(R(X), np.float64)
    H1 = np.asarray(R(np.maximum(Xr @ W1 + np.asarray(p["b1"], np.float64), 0.0)), np.float64)
    H2 = np.asarray(R(np.maximum(H1 @ W2 + np.asarray(p["b2"], np.float64), 0.0)), np.float64)
    z = H2 @ np.asarray(p["w3"], np.float64) + float(p["b3"][0])
    prob = 1.0 / (1.0 + np.exp(-z))
    return prob, (Xr, W1, W2, H1, H2, z)


def loss_and_grads(X, y, p, rounding=True):
    R = bf16 if rounding else (lambda a: np.asarray(a, np.float64))   # noqa: E731
    prob, (Xr, W1, W2, H1, H2, z) = forward(X, p, rounding)
    n = len(y)
    y = np.asarray(y, np.float64)
    loss = float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (prob - y) / n
    w3 = np.asarray(p["w3"], np.float64)
    dZ2 = np.asarray(R(np.where(H2 > 0, dz[:, None] * w3[None, :], 0.0)), np.float64)
    g = {"w3": H2.T @ dz, "b3": np.array([dz.sum()]),
         "W2": H1.T @ dZ2, "b2": dZ2.sum(0)}
    dZ1 = np.asarray(R(np.where(H1 > 0, dZ2 @ W2.T, 0.0)), np.float64)
    g["W1"] = Xr.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return loss, g


def adam(w, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step in float32 (returns new w, m, v)."""
    w, m, v, g = (np.asarray(a, np.float32) for a in (w, m, v, g))
    f = np.float32
    m = f(b1) * m + f(1 - b1) * g
    v = f(b2) * v + f(1 - b2) * g * g
    c1 = f(1 - b1 ** step)
    c2 = f(1 - b2 ** step)
    w = w - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w, m, v
