"""NumPy restatement of the learned densification proposal network (16 -> 64 -> 64 -> 3).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the MFMA kernels. The product
never imports it.

The network has no counterpart in the reference (it is a new component). The restatement
mirrors the device kernels' quantisation points: features, weights, hidden activations and the
backward's dZ tensors are rounded to BF16 (round to nearest even) where the device feeds them to
the matrix cores; every sum is taken in float64 here (the device sums in FP32, in its own order),
so comparisons use a tolerance. With quantize=False the same code is the exact float64 network,
which the tests pin against central finite differences.

Parameter layout (float32[5443]): W1[64][16], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].
"""
from __future__ import annotations

import numpy as np

IN = 16
HIDDEN = 64
CLASSES = 3
SHAPES = (("W1", (HIDDEN, IN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("W3", (CLASSES, HIDDEN)), ("b3", (CLASSES,)))
PARAMETERS = sum(int(np.prod(shape)) for _, shape in SHAPES)
BF16_WEIGHTS = HIDDEN * IN + HIDDEN * HIDDEN + 32 * HIDDEN


def bf16(x) -> np.ndarray:
    """Rounds float32 values to bfloat16 (nearest even), returned as float32. NaN stays NaN."""
    x = np.asarray(x, dtype=np.float32)
    bits = x.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), np.float32(np.nan), out).astype(np.float32)


def bf16_bits(x) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of float32 values, nearest even."""
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def unpack(params: np.ndarray) -> dict:
    params = np.asarray(params, dtype=np.float32)
    assert params.shape == (PARAMETERS,)
    out, offset = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = params[offset:offset + size].reshape(shape)
        offset += size
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], np.float32).reshape(-1) for name, _ in SHAPES])


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[1])
            parts[name] = rng.uniform(-bound, bound, shape).astype(np.float32)
        else:
            parts[name] = np.zeros(shape, np.float32)
    return pack(parts)


def bf16_weight_block(params: np.ndarray) -> np.ndarray:
    """The uint16[7168] BF16 copy the device reads: W1, W2, W3 padded to 32 rows."""
    p = unpack(params)
    w3 = np.zeros((32, HIDDEN), np.float32)
    w3[:CLASSES] = p["W3"]
    return np.concatenate([bf16_bits(p["W1"]).reshape(-1), bf16_bits(p["W2"]).reshape(-1),
                           bf16_bits(w3).reshape(-1)])


def _q(x, quantize: bool) -> np.ndarray:
    return bf16(np.asarray(x, np.float32)).astype(np.float64) if quantize else np.asarray(x, np.float64)


def forward(features: np.ndarray, params: np.ndarray, quantize: bool = True) -> dict:
    """Logits and the intermediates of the backward pass."""
    p = unpack(params)
    x = _q(features, quantize)
    w1, w2, w3 = (_q(p[k], quantize) for k in ("W1", "W2", "W3"))
    b1, b2, b3 = (np.asarray(p[k], np.float64) for k in ("b1", "b2", "b3"))
    z1 = x @ w1.T + b1
    a1 = _q(np.maximum(z1, 0.0).astype(np.float32), quantize) if quantize else np.maximum(z1, 0.0)
    z2 = a1 @ w2.T + b2
    a2 = _q(np.maximum(z2, 0.0).astype(np.float32), quantize) if quantize else np.maximum(z2, 0.0)
    z3 = a2 @ w3.T + b3
    return {"x": x, "w1": w1, "w2": w2, "w3": w3, "z1": z1, "a1": a1, "z2": z2, "a2": a2, "logits": z3}


def loss_and_gradients(features, labels, params, sample_weights=None, quantize: bool = True,
                       magnitudes: bool = False):
    """Weighted mean softmax cross-entropy (weights w_i / n) and its parameter gradients. With
    magnitudes=True also returns, per parameter, the sum of the absolute values of the per-sample
    terms that make up its gradient (the scale of its rounding error in any summation order)."""
    features = np.asarray(features, np.float32)
    n = features.shape[0]
    labels = np.asarray(labels, np.int64)
    w = np.ones(n) if sample_weights is None else np.asarray(sample_weights, np.float64)
    w = w * (np.float64(np.float32(1.0 / n)) if quantize else 1.0 / n)
    f = forward(features, params, quantize)
    z3 = f["logits"]
    top = z3.max(axis=1, keepdims=True)
    e = np.exp(z3 - top)
    s = e.sum(axis=1, keepdims=True)
    loss = float(np.sum(w * (top[:, 0] + np.log(s[:, 0]) - z3[np.arange(n), labels])))
    onehot = np.zeros_like(z3)
    onehot[np.arange(n), labels] = 1.0
    dz3 = w[:, None] * (e / s - onehot)
    if quantize:
        dz3 = _q(dz3.astype(np.float32), True)
    dw3 = dz3.T @ f["a2"]
    db3 = dz3.sum(axis=0)
    da2 = dz3 @ f["w3"]
    dz2 = np.where(f["z2"] > 0.0, da2, 0.0)
    if quantize:
        dz2 = _q(dz2.astype(np.float32), True)
    dw2 = dz2.T @ f["a1"]
    db2 = dz2.sum(axis=0)
    da1 = dz2 @ f["w2"]
    dz1 = np.where(f["z1"] > 0.0, da1, 0.0)
    if quantize:
        dz1 = _q(dz1.astype(np.float32), True)
    dw1 = dz1.T @ f["x"]
    db1 = dz1.sum(axis=0)
    gradients = np.concatenate([dw1.reshape(-1), db1, dw2.reshape(-1), db2, dw3.reshape(-1), db3])
    if not magnitudes:
        return loss, gradients
    a1, a2, x = np.abs(f["a1"]), np.abs(f["a2"]), np.abs(f["x"])
    m3, m2, m1 = np.abs(dz3), np.abs(dz2), np.abs(dz1)
    scale = np.concatenate([(m1.T @ x).reshape(-1), m1.sum(axis=0), (m2.T @ a1).reshape(-1), m2.sum(axis=0),
                            (m3.T @ a2).reshape(-1), m3.sum(axis=0)])
    return loss, gradients, scale


def adam_f32(params, first, second, gradients, learning_rate, step):
    """Adam with beta 0.9 / 0.999, eps 1e-8, in float32 (bias corrections from float64)."""
    g = np.asarray(gradients, np.float32)
    c1 = np.float32(1.0 / (1.0 - 0.9 ** step))
    c2 = np.float32(1.0 / (1.0 - 0.999 ** step))
    m = np.float32(0.9) * first + np.float32(0.1) * g
    v = np.float32(0.999) * second + np.float32(0.001) * g * g
    update = np.float32(learning_rate) * (m * c1) / (np.sqrt(v * c2) + np.float32(1e-8))
    return (params - update).astype(np.float32), m.astype(np.float32), v.astype(np.float32)
