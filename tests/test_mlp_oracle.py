"""Pins of the NumPy restatement of the densification proposal network (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import mlp


def test_bf16_rounding_matches_torch():
    # torch's float32 -> bfloat16 conversion rounds to nearest even.
    rng = np.random.default_rng(0)
    values = np.concatenate([
        rng.standard_normal(4096).astype(np.float32) * 10.0,
        np.array([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0078125, 3.0e38, -3.0e38, 1e-40, np.inf, -np.inf],
                 np.float32),
        # exact halfway cases between two bf16 values: ties to even both ways
        (np.arange(1, 200, dtype=np.uint32) * 0x10000 + 0x8000).view(np.float32),
    ])
    expected = torch.from_numpy(values).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mlp.bf16(values), expected)
    assert np.isnan(mlp.bf16(np.array([np.nan], np.float32)))[0]


def test_pack_layout():
    params = np.arange(mlp.PARAMETERS, dtype=np.float32)
    parts = mlp.unpack(params)
    assert mlp.PARAMETERS == 5443
    assert parts["W1"][1, 0] == 16 and parts["b1"][0] == 1024 and parts["W2"][0, 1] == 1089
    assert parts["b3"][-1] == 5442
    np.testing.assert_array_equal(mlp.pack(parts), params)
    block = mlp.bf16_weight_block(params)
    assert block.shape == (mlp.BF16_WEIGHTS,) and block.dtype == np.uint16
    assert np.all(block[64 * 16 + 64 * 64 + 3 * 64:] == 0)


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    features = rng.standard_normal((n, mlp.IN)).astype(np.float32)
    labels = rng.integers(0, 3, n)
    weights = rng.uniform(0.5, 2.0, n).astype(np.float32)
    params = mlp.init_params(seed)
    parts = mlp.unpack(params)
    parts["b1"] = rng.uniform(-0.3, 0.3, mlp.HIDDEN).astype(np.float32)
    parts["b2"] = rng.uniform(-0.3, 0.3, mlp.HIDDEN).astype(np.float32)
    parts["b3"] = rng.uniform(-0.3, 0.3, mlp.CLASSES).astype(np.float32)
    return features, labels, weights, mlp.pack(parts)


def test_forward_closed_form():
    # W1 copies the first 3 inputs (and their negations) into 6 hidden units; W2 passes them on;
    # W3 recombines x_c = relu(x_c) - relu(-x_c). With BF16-exact inputs the logits equal them.
    parts = {name: np.zeros(shape, np.float32) for name, shape in mlp.SHAPES}
    for c in range(3):
        parts["W1"][2 * c, c] = 1.0
        parts["W1"][2 * c + 1, c] = -1.0
        parts["W2"][2 * c, 2 * c] = 1.0
        parts["W2"][2 * c + 1, 2 * c + 1] = 1.0
        parts["W3"][c, 2 * c] = 1.0
        parts["W3"][c, 2 * c + 1] = -1.0
    parts["b3"][:] = [0.5, -0.25, 2.0]
    x = np.zeros((5, mlp.IN), np.float32)
    x[:, :3] = [[1, -2, 3], [0.5, 0.25, -0.125], [0, 0, 0], [-7, 8, 1.5], [100, -100, 0.75]]
    logits = mlp.forward(x, mlp.pack(parts))["logits"]
    np.testing.assert_array_equal(logits, x[:, :3] + parts["b3"])


@pytest.mark.parametrize("use_weights", [False, True])
def test_gradients_match_finite_differences(use_weights):
    features, labels, weights, params = _problem(40, 3)
    sample_weights = weights if use_weights else None
    _, gradients = mlp.loss_and_gradients(features, labels, params, sample_weights, quantize=False)
    rng = np.random.default_rng(7)
    # every bias plus a random subset of weights in each layer
    picks = list(range(1024, 1088)) + list(range(5184, 5248)) + list(range(5440, 5443))
    picks += list(rng.choice(1024, 40, replace=False)) + list(1088 + rng.choice(4096, 60, replace=False))
    picks += list(5248 + rng.choice(192, 40, replace=False))
    h = 1e-4
    base = params.astype(np.float64)
    for index in picks:
        plus, minus = base.copy(), base.copy()
        plus[index] += h
        minus[index] -= h
        loss_plus = _loss64(features, labels, plus, sample_weights)
        loss_minus = _loss64(features, labels, minus, sample_weights)
        numeric = (loss_plus - loss_minus) / (2 * h)
        assert abs(numeric - gradients[index]) <= 1e-6 + 1e-5 * abs(numeric), (index, numeric, gradients[index])


def _loss64(features, labels, params64, sample_weights):
    # float64 parameters: the float32 packing in forward() would round the perturbation away
    n = features.shape[0]
    sizes = [int(np.prod(s)) for _, s in mlp.SHAPES]
    offsets = np.cumsum([0] + sizes)
    w1, b1, w2, b2, w3, b3 = (params64[offsets[i]:offsets[i + 1]].reshape(mlp.SHAPES[i][1]) for i in range(6))
    x = features.astype(np.float64)
    a1 = np.maximum(x @ w1.T + b1, 0)
    a2 = np.maximum(a1 @ w2.T + b2, 0)
    z = a2 @ w3.T + b3
    top = z.max(axis=1)
    lse = top + np.log(np.exp(z - top[:, None]).sum(axis=1))
    w = np.ones(n) if sample_weights is None else sample_weights.astype(np.float64)
    return float(np.sum(w / n * (lse - z[np.arange(n), labels])))


def test_quantized_gradients_close_to_exact():
    # BF16 rounding perturbs the gradients by about the BF16 precision (2^-8), not more.
    features, labels, weights, params = _problem(256, 5)
    _, exact = mlp.loss_and_gradients(features, labels, params, weights, quantize=False)
    _, quantized = mlp.loss_and_gradients(features, labels, params, weights, quantize=True)
    scale = np.abs(exact).max()
    assert np.abs(exact - quantized).max() < 0.03 * scale


def test_adam_f32_first_step_is_sign_step():
    # At step 1 the bias-corrected update is lr * g / (|g| + eps) = lr * sign(g) for |g| >> eps.
    rng = np.random.default_rng(1)
    g = rng.standard_normal(mlp.PARAMETERS).astype(np.float32)
    params = rng.standard_normal(mlp.PARAMETERS).astype(np.float32)
    zeros = np.zeros_like(params)
    new, m, v = mlp.adam_f32(params, zeros, zeros, g, 1e-3, 1)
    # params are O(1): their float32 spacing (1.2e-7) bounds the recovered step
    np.testing.assert_allclose(params - new, 1e-3 * np.sign(g), rtol=2e-4, atol=3e-7)
    np.testing.assert_array_equal(m, np.float32(0.1) * g)
