"""FP64 restatement of the BFS variant predictor (11 -> 32 -> 32 -> K MLP).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
inference, gradient and training kernels -- never by the product path.

The reference has no network (its switch models are decision trees built
offline, benchmark-analysis/model-src); the function pinned here is plain
mathematics: a two-hidden-layer ReLU perceptron on normalised inputs and
the mean squared error.  Gradients are checked by central finite
differences of :func:`loss`, which depend on :func:`forward` only, not on
any restatement of the backward pass.
"""
from __future__ import annotations

import numpy as np

F, H = 11, 32  # features (BW_BFS_FEATURES), hidden width


def layout(K: int) -> dict:
    o = {"mean": 0, "std": F, "W1": 2 * F}
    o["b1"] = o["W1"] + H * F
    o["W2"] = o["b1"] + H
    o["b2"] = o["W2"] + H * H
    o["W3"] = o["b2"] + H
    o["b3"] = o["W3"] + K * H
    o["end"] = o["b3"] + K
    return o


def unpack(w: np.ndarray, K: int):
    L = layout(K)
    w = np.asarray(w, dtype=np.float64)
    return (w[L["mean"]:L["std"]], w[L["std"]:L["W1"]], w[L["W1"]:L["b1"]].reshape(H, F), w[L["b1"]:L["W2"]],
            w[L["W2"]:L["b2"]].reshape(H, H), w[L["b2"]:L["W3"]], w[L["W3"]:L["b3"]].reshape(K, H),
            w[L["b3"]:L["end"]])


def forward(w: np.ndarray, K: int, X: np.ndarray) -> np.ndarray:
    mean, std, W1, b1, W2, b2, W3, b3 = unpack(w, K)
    x = (np.asarray(X, dtype=np.float64) - mean) / std
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


def loss(w: np.ndarray, K: int, X: np.ndarray, Y: np.ndarray) -> float:
    r = forward(w, K, X) - np.asarray(Y, dtype=np.float64)
    return float(np.mean(r * r))


def fd_gradient(w: np.ndarray, K: int, X: np.ndarray, Y: np.ndarray, h: float = 1e-6,
                indices=None) -> np.ndarray:
    """Central differences of :func:`loss` w.r.t. the packed parameters
    (only `indices` if given; the others are left 0)."""
    w = np.asarray(w, dtype=np.float64).copy()
    g = np.zeros_like(w)
    idx = range(2 * F, w.size) if indices is None else indices
    for i in idx:
        old = w[i]
        w[i] = old + h
        lp = loss(w, K, X, Y)
        w[i] = old - h
        lm = loss(w, K, X, Y)
        w[i] = old
        g[i] = (lp - lm) / (2 * h)
    return g
