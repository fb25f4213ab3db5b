"""TEST INFRASTRUCTURE — CPU restatement of the learned seed scorer's arithmetic.

Forward, loss and gradients of the 10 -> 32 -> 32 -> 1 MLP with the BF16 roundings at the points the GPU kernels use (allen_amd/csrc/velo_scorer.hip). Sums are in float64; the GPU accumulates in FP32. The scorer has no counterpart in the reference.

**Pins.**
- The matrix products are checked against numpy's own matmul on exactly representable inputs.
- The gradients are checked against central finite differences of `loss`, in float64 and without the bf16 rounding.

Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 10, 32
W1, B1, W2, B2, W3, B3 = 0, 320, 352, 1376, 1408, 1440
N_PARAMS = 1441


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16, returned as float64 (as the GPU's to_bf16)."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(N_IN, N_H), p[B1:W2], p[W2:B2].reshape(N_H, N_H), p[B2:W3], p[W3:B3], p[B3])


def forward(x: np.ndarray, p: np.ndarray, rounded: bool = True) -> dict:
    w1, b1, w2, b2, w3, b3 = unpack(p)
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    xb = r(x)
    z1 = xb @ r(w1) + b1
    a1b = r(np.maximum(z1, 0.0).astype(np.float32)) if rounded else np.maximum(z1, 0.0)
    z2 = a1b @ r(w2) + b2
    a2 = np.maximum(z2, 0.0)
    w3b = r(w3)
    z3 = a2 @ w3b + b3
    return {"xb": xb, "z1": z1, "a1b": a1b, "z2": z2, "a2": a2, "w3b": w3b, "z3": z3, "w1b": r(w1), "w2b": r(w2)}


def loss(x, y, p, rounded: bool = True) -> float:
    z3 = forward(x, p, rounded)["z3"]
    return float(np.sum(np.maximum(z3, 0) - z3 * y + np.log1p(np.exp(-np.abs(z3)))))


def gradient(x: np.ndarray, y: np.ndarray, p: np.ndarray, rounded: bool = True) -> np.ndarray:
    """Gradient of the mean loss over the rows."""
    f = forward(x, p, rounded)
    r = bf16 if rounded else (lambda a: np.asarray(a, dtype=np.float64))
    n = x.shape[0]
    sig = 1.0 / (1.0 + np.exp(-f["z3"]))
    dz3 = (sig - y) / n
    dz2 = np.where(f["z2"] > 0, dz3[:, None] * f["w3b"][None, :], 0.0)
    dz2b = r(dz2.astype(np.float32)) if rounded else dz2
    g = np.zeros(N_PARAMS)
    g[W2:B2] = (f["a1b"].T @ dz2b).ravel()
    g[B2:W3] = dz2.sum(axis=0)
    g[W3:B3] = dz3 @ f["a2"]
    g[B3] = dz3.sum()
    da1 = dz2b @ f["w2b"].T
    dz1 = np.where(f["z1"] > 0, da1, 0.0)
    dz1b = r(dz1.astype(np.float32)) if rounded else dz1
    g[W1:B1] = (f["xb"].T @ dz1b).ravel()
    g[B1:W2] = dz1.sum(axis=0)
    return g
