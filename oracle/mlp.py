"""FP64 restatement of the radius-predictor MLP (TEST INFRASTRUCTURE ONLY).

12 -> 64 -> 64 -> 1 ReLU network, parameters flat in the library's layout:
W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3.  Loss = mean (f(x)-t)^2.
Used to check the GPU forward and backward kernels; pinned by central finite
differences in tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np

F_IN, H = 12, 64
SIZES = [("W1", (H, F_IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, bf16_inference=False):
    w = unpack(p)
    X = np.asarray(X, np.float64)
    if bf16_inference:
        w = {k: bf16(v) for k, v in w.items()}
        X = bf16(X)
    h1 = np.maximum(X @ w["W1"].T + w["b1"], 0.0)
    if bf16_inference:
        h1 = bf16(h1)
    h2 = np.maximum(h1 @ w["W2"].T + w["b2"], 0.0)
    if bf16_inference:
        h2 = bf16(h2)
    return h2 @ w["W3"] + w["b3"][0]


def loss(p, X, T):
    y = forward(p, X)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def gradient(p, X, T):
    w = unpack(p)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    a1 = X @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w["W3"] + w["b3"][0]
    dy = 2.0 * (y - T) / n
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()])}
    d2 = np.outer(dy, w["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SIZES]), float(np.mean((y - T) ** 2))
