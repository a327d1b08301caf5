"""FP64 numpy restatement of the learned-warm-start MLP (forward, backward,
Adam, BF16-weight inference).

TEST INFRASTRUCTURE ONLY: the product path runs the libadx kernels
(adx_mlp_*); this module checks them.  The network has no counterpart in the
reference repository (its Q initialiser is randomized SVD + ALS,
adamixture/model/als_gpu.py:45-127, whose Q map at :23-36 the inference
epilogue reproduces).

Pinned in tests/test_oracle.py by central finite differences of ``loss``
(every parameter block, both signs of the ReLU), by the closed form of the
softmax cross-entropy gradient on a one-layer case, and by the textbook Adam
update (Kingma & Ba 2015, Algorithm 1).
"""
from __future__ import annotations

import numpy as np

D0, D1, D2, D3 = 64, 256, 256, 16
SHAPES = [("W1", (D0, D1)), ("b1", (D1,)), ("W2", (D1, D2)), ("b2", (D2,)), ("W3", (D2, D3)), ("b3", (D3,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(w: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = w[o:o + n].reshape(shp)
        o += n
    return out


def forward(w: np.ndarray, X: np.ndarray, kvec: np.ndarray):
    """Returns (probs, cache).  probs[r, c] = 0 for c >= kvec[r]."""
    p = unflatten(np.asarray(w, np.float64))
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ p["W1"] + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["W2"] + p["b2"], 0.0)
    z3 = h2 @ p["W3"] + p["b3"]
    z = z3 + np.log(np.maximum(X[:, :D3], 1e-5))
    mask = np.arange(D3)[None, :] < np.asarray(kvec)[:, None]
    z = np.where(mask, z, -np.inf)
    z = z - z.max(axis=1, keepdims=True)
    e = np.where(mask, np.exp(z), 0.0)
    probs = e / e.sum(axis=1, keepdims=True)
    return probs, (X, h1, h2, mask, p)


def loss(w, X, kvec, T) -> float:
    """Mean over rows of -sum_c t log p (masked columns excluded)."""
    probs, (_, _, _, mask, _) = forward(w, X, kvec)
    lp = np.where(mask, np.log(np.where(mask, probs, 1.0)), 0.0)
    return float(-(np.asarray(T, np.float64) * lp).sum() / X.shape[0])


def backward(w, X, kvec, T):
    """Gradient of ``loss`` w.r.t. the flat parameters (same layout)."""
    probs, (X, h1, h2, mask, p) = forward(w, X, kvec)
    rows = X.shape[0]
    dz = np.where(mask, probs - np.asarray(T, np.float64), 0.0) / rows
    g = {}
    g["W3"] = h2.T @ dz
    g["b3"] = dz.sum(0)
    dh2 = (dz @ p["W3"].T) * (h2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ p["W2"].T) * (h1 > 0)
    g["W1"] = X.T @ dh1
    g["b1"] = dh1.sum(0)
    return np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (returns new w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


def _bf16(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def infer_bf16(w, X, K: int) -> np.ndarray:
    """Inference with BF16 weights and activations (products exact, FP64
    sums), then the Q map: clamp [1e-5, 1-1e-5] and renormalise."""
    p = unflatten(np.asarray(w, np.float64))
    Xb = _bf16(X)
    h1 = _bf16(np.maximum(Xb @ _bf16(p["W1"]) + p["b1"].astype(np.float32), 0.0))
    h2 = _bf16(np.maximum(h1 @ _bf16(p["W2"]) + p["b2"].astype(np.float32), 0.0))
    z3 = h2 @ _bf16(p["W3"]) + p["b3"].astype(np.float32)
    z = z3[:, :K] + np.log(np.maximum(np.asarray(X, np.float64)[:, :K], 1e-5))
    z -= z.max(axis=1, keepdims=True)
    q = np.exp(z)
    q /= q.sum(axis=1, keepdims=True)
    q = np.clip(q, 1e-5, 1 - 1e-5)
    return q / q.sum(axis=1, keepdims=True)

