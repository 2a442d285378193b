"""CPU reference of the learned part's MLP (TEST INFRASTRUCTURE ONLY).

80 -> 128 (ReLU) -> 128 (ReLU) -> 2, loss = mean BCE-with-logits(z0, y)
+ lambda * mean (z1 - c)^2.  Every matrix product rounds both operands to
bfloat16 (round to nearest even), as the device GEMM stages them, and
accumulates in float64.  The reference FARM has no learned component, so
this module is pinned by itself: its analytic gradient is checked against
central finite differences of its own float64 loss (tests/test_mlp_ref.py).
"""
from __future__ import annotations

import numpy as np

NFEAT, HID, OUT = 80, 128, 2
SHAPES = [("W1", (NFEAT, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)),
          ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, rounding):
    if rounding:
        return bf16(a) @ bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(p: np.ndarray, X: np.ndarray, rounding: bool = True):
    P = unpack(np.asarray(p, np.float64))
    z1 = _mm(X, P["W1"], rounding) + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = _mm(h1, P["W2"], rounding) + P["b2"]
    h2 = np.maximum(z2, 0)
    z3 = _mm(h2, P["W3"], rounding) + P["b3"]
    return z3, (h1, h2)


def loss(p, X, y, c, lam, rounding: bool = True) -> float:
    z, _ = forward(p, X, rounding)
    z0, z1 = z[:, 0], z[:, 1]
    bce = np.maximum(z0, 0) - z0 * y + np.log1p(np.exp(-np.abs(z0)))
    return float(bce.mean() + lam * ((z1 - c) ** 2).mean())


def loss_grad(p, X, y, c, lam, rounding: bool = True):
    """Loss and its gradient (flat, NPARAMS) by backpropagation."""
    P = unpack(np.asarray(p, np.float64))
    M = len(X)
    z, (h1, h2) = forward(p, X, rounding)
    z0, z1 = z[:, 0], z[:, 1]
    s = 1.0 / (1.0 + np.exp(-z0))
    d3 = np.stack([(s - y) / M, 2 * lam * (z1 - c) / M], 1)
    g = {}
    g["W3"] = _mm(h2.T, d3, rounding)
    g["b3"] = d3.sum(0)
    dh2 = _mm(d3, P["W3"].T, rounding) * (h2 > 0)
    g["W2"] = _mm(h1.T, dh2, rounding)
    g["b2"] = dh2.sum(0)
    dh1 = _mm(dh2, P["W2"].T, rounding) * (h1 > 0)
    g["W1"] = _mm(np.asarray(X).T, dh1, rounding)
    g["b1"] = dh1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    bce = np.maximum(z0, 0) - z0 * y + np.log1p(np.exp(-np.abs(z0)))
    return float(bce.mean() + lam * ((z1 - c) ** 2).mean()), flat
