"""CPU oracle of the tile-cost MLP (TEST INFRASTRUCTURE ONLY: loaded by tests/ and never by the
product path).

The network has no counterpart in the reference: it is the learned tile-queue order of the
batched renderer (24 features -> 64 -> 64 -> 1, ReLU, BF16 weights and activations, FP32
accumulation). This module restates the HIP kernels' arithmetic:

* forward: every weight, bias and layer input is rounded to BF16 (round to nearest even);
  each output starts from its bias and adds w * x for k = 0, 1, ... in float32 (the products
  of two BF16 values are exact in float32, so the only rounding is the sequential sum);
* gradients of the mean squared error, with the BF16 rounding of activations treated as the
  identity (straight-through), computed in float64 from the same rounded values.

Parameter layout (one float32 vector): W1[64][24], b1[64], W2[64][64], b2[64], w3[64], b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 24, 64
SIZES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("w3", (N_HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[off:off + n].reshape(shape)
        off += n
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even) -> float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (rounded & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _layer(w: np.ndarray, b: np.ndarray, x: np.ndarray) -> np.ndarray:
    """float32 sequential accumulation: acc = b; acc += w[:, k] * x[:, k] for k in order."""
    acc = np.broadcast_to(b.astype(np.float32), (x.shape[0], w.shape[0])).copy()
    for k in range(w.shape[1]):
        acc = (acc + np.outer(x[:, k], w[:, k]).astype(np.float32)).astype(np.float32)
    return acc


def forward(p: np.ndarray, X: np.ndarray, keep: bool = False, quantize: bool = True):
    """quantize=False evaluates the same network in float64 without BF16 rounding (used to pin
    the gradient formulas against finite differences)."""
    if quantize:
        q = {k: bf16(v) for k, v in unpack(np.asarray(p, dtype=np.float32)).items()}
        xb = bf16(X.astype(np.float32))
        z1 = _layer(q["W1"], q["b1"], xb)
        h1 = bf16(np.maximum(z1, 0.0).astype(np.float32))
        z2 = _layer(q["W2"], q["b2"], h1)
        h2 = bf16(np.maximum(z2, 0.0).astype(np.float32))
        y = _layer(q["w3"][None, :], q["b3"], h2)[:, 0]
    else:
        q = {k: v.astype(np.float64) for k, v in unpack(np.asarray(p, dtype=np.float64)).items()}
        xb = X.astype(np.float64)
        z1 = xb @ q["W1"].T + q["b1"]
        h1 = np.maximum(z1, 0.0)
        z2 = h1 @ q["W2"].T + q["b2"]
        h2 = np.maximum(z2, 0.0)
        y = h2 @ q["w3"] + q["b3"][0]
    if keep:
        return y, dict(q=q, xb=xb, z1=z1, h1=h1, z2=z2, h2=h2)
    return y


def gradients(p: np.ndarray, X: np.ndarray, t: np.ndarray, quantize: bool = True) -> np.ndarray:
    """d mean((y - t)^2) / d params, float64, same layout as the parameter vector."""
    y, c = forward(p, X, keep=True, quantize=quantize)
    q = {k: v.astype(np.float64) for k, v in c["q"].items()}
    B = X.shape[0]
    dy = 2.0 * (y.astype(np.float64) - t.astype(np.float64)) / B
    h2, h1, xb = c["h2"].astype(np.float64), c["h1"].astype(np.float64), c["xb"].astype(np.float64)
    dw3 = dy @ h2
    db3 = np.array([dy.sum()])
    dz2 = np.outer(dy, q["w3"]) * (c["z2"] > 0)
    dW2 = dz2.T @ h1
    db2 = dz2.sum(axis=0)
    dz1 = (dz2 @ q["W2"]) * (c["z1"] > 0)
    dW1 = dz1.T @ xb
    db1 = dz1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, db3])


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (the same initialisation the library uses is not
    required; tests pass parameters explicitly)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("b"):
            parts.append(np.zeros(shape, np.float32).ravel())
        else:
            fan_in = shape[-1]
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, size=shape).astype(np.float32).ravel())
    return np.concatenate(parts).astype(np.float32)
