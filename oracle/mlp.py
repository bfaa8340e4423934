"""CPU oracle of the per-node MLP (TEST INFRASTRUCTURE ONLY): forward, backward and Adam in
NumPy with the device's precision recipe -- GEMM operands rounded to bf16 (round to nearest even),
products accumulated in float64 here (the device accumulates in FP32), tanh hidden layers, linear
output, MSE loss.  No reference counterpart (the reference has no learned predictor); the pins are
mathematical: backward agrees with central finite differences of the float64 forward, and the
device agrees with this oracle within bf16 tolerances.
"""

from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 precision (nearest even) and return them as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(W, b, X, *, round_bf16=True):
    """Activations list [X, Y1, ..., YL] (float64)."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [np.asarray(X, dtype=np.float64)]
    for li, (w, bb) in enumerate(zip(W, b)):
        z = q(acts[-1]) @ q(w) + bb
        acts.append(np.tanh(z) if li < len(W) - 1 else z)
    return acts


def backward(W, acts, dOut, *, round_bf16=True):
    """Gradients (gW, gb) of sum(dOut * out)."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    gW = [None] * len(W)
    gb = [None] * len(W)
    dZ = np.asarray(dOut, dtype=np.float64)
    for li in range(len(W) - 1, -1, -1):
        gW[li] = q(acts[li]).T @ q(dZ)
        gb[li] = dZ.sum(axis=0)
        if li == 0:
            break
        dX = q(dZ) @ q(W[li]).T
        dZ = dX * (1.0 - acts[li] ** 2)
    return gW, gb


def mse(pred, target, mask=None):
    m = np.ones_like(pred) if mask is None else mask
    cnt = m.sum()
    d = (pred - target) * m
    return float((d * d).sum() / cnt), 2.0 * d / cnt


def adam(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
