"""MFMA proposal network (16 -> 64 -> 64 -> 3, BF16 operands, FP32 accumulation) against the NumPy
restatement in oracle/mlp.py.

Tolerances: the device and the restatement round the same tensors to BF16, but the device sums
in FP32 in the matrix cores' order and the restatement in float64. An FP32 value that lands next
to a BF16 rounding midpoint can therefore round one BF16 step (2^-8 relative) differently, and the
step propagates through the later layers. Logits are compared at 1e-2 of the logits' scale per
sample set. Gradients are compared per element against the sum of the absolute values of the
per-sample terms that form them (5e-5 of that magnitude), per layer normwise (1e-5), and by the
median error over magnitude (1e-6); a transposed or misplaced operand fragment, or a BF16
rounding on the other side of a midpoint (up to 2^-7 of one term), would not meet them.
"""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _problem(n, seed, bias=0.3):
    rng = np.random.default_rng(seed)
    features = (rng.standard_normal((n, mlp.IN)) * rng.uniform(0.2, 3.0, mlp.IN)).astype(np.float32)
    labels = rng.integers(0, 3, n).astype(np.int32)
    weights = rng.uniform(0.5, 2.0, n).astype(np.float32)
    parts = mlp.unpack(mlp.init_params(seed))
    for name in ("b1", "b2", "b3"):
        parts[name] = rng.uniform(-bias, bias, parts[name].shape).astype(np.float32)
    return features, labels, weights, mlp.pack(parts)


def test_bf16_weight_copy_matches_round_to_nearest_even(gsn, context):
    _, _, _, params = _problem(1, 11)
    params[:7] = [1.00390625, 1.005859375, -1.00390625, 3.0e38, 1e-40, 0.0, -0.0]
    net = gsn.ProposalNetwork(context, params)
    context.synchronize()
    np.testing.assert_array_equal(net.weights.numpy(), mlp.bf16_weight_block(params))


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 40000])
def test_forward_matches_oracle(gsn, context, n):
    features, _, _, params = _problem(n, n)
    net = gsn.ProposalNetwork(context, params)
    x = gsn.DeviceBuffer.from_numpy(features)
    logits = gsn.DeviceBuffer((n, 3), np.float32)
    net.forward(x, n, logits)
    context.synchronize()
    actual = logits.numpy()
    expected = mlp.forward(features, params)["logits"]
    error = np.abs(actual - expected)
    scale = np.abs(expected).max()
    assert error.max() <= 1e-2 * scale, (error.max(), scale)
    assert np.median(error) <= 1e-3 * scale


def test_forward_closed_form(gsn, context):
    # W1/W2/W3 route inputs 0..2 through relu(x) - relu(-x): the logits equal x[:, :3] + b3 exactly
    # for BF16-exact inputs, whatever the accumulation order.
    parts = {name: np.zeros(shape, np.float32) for name, shape in mlp.SHAPES}
    for c in range(3):
        parts["W1"][2 * c, c] = 1.0
        parts["W1"][2 * c + 1, c] = -1.0
        parts["W2"][2 * c, 2 * c] = 1.0
        parts["W2"][2 * c + 1, 2 * c + 1] = 1.0
        parts["W3"][c, 2 * c] = 1.0
        parts["W3"][c, 2 * c + 1] = -1.0
    parts["b3"][:] = [0.5, -0.25, 2.0]
    rng = np.random.default_rng(2)
    n = 77
    features = mlp.bf16(rng.standard_normal((n, mlp.IN)).astype(np.float32) * 4)
    net = gsn.ProposalNetwork(context, mlp.pack(parts))
    logits = gsn.DeviceBuffer((n, 3), np.float32)
    net.forward(gsn.DeviceBuffer.from_numpy(features), n, logits)
    context.synchronize()
    np.testing.assert_array_equal(logits.numpy(), features[:, :3] + parts["b3"])


def _layer_slices():
    offset, out = 0, []
    for name, shape in mlp.SHAPES:
        size = int(np.prod(shape))
        out.append((name, slice(offset, offset + size)))
        offset += size
    return out


@pytest.mark.parametrize("n,weighted", [(1, False), (45, True), (4096, True), (100000, False)])
def test_gradients_and_loss_match_oracle(gsn, context, n, weighted):
    features, labels, weights, params = _problem(n, 100 + n)
    sample_weights = weights if weighted else None
    net = gsn.ProposalNetwork(context, params)
    gradients = gsn.DeviceBuffer((mlp.PARAMETERS,), np.float32)
    loss = net.train_step(gsn.DeviceBuffer.from_numpy(features), gsn.DeviceBuffer.from_numpy(labels), n, 0.0,
                          gsn.DeviceBuffer.from_numpy(sample_weights) if weighted else None, gradients, want_loss=True)
    actual = gradients.numpy()
    expected_loss, expected, magnitude = mlp.loss_and_gradients(features, labels, params, sample_weights,
                                                                magnitudes=True)
    assert abs(loss - expected_loss) <= 1e-3 * abs(expected_loss) + 1e-6
    for name, part in _layer_slices():
        error = np.abs(actual[part] - expected[part])
        normwise = np.linalg.norm(actual[part] - expected[part]) / np.linalg.norm(expected[part])
        relative = error / np.maximum(magnitude[part], 1e-30)
        print(f"n {n} {name}: normwise {normwise:.2e}, max error / magnitude {relative.max():.2e}, "
              f"median {np.median(relative):.2e}", flush=True)
        # Measured on these seeds: at most 5.7e-6 of the magnitude per element, 1.6e-6 normwise,
        # medians at most 2.3e-7 (FP32 summation order only: no BF16 rounding landed on the other
        # side of a midpoint, which would move one term by up to 2^-7 of itself).
        assert np.all(error <= 5e-5 * magnitude[part] + 1e-12), (name, relative.max())
        assert normwise <= 1e-5, (name, normwise)
        assert np.median(relative) <= 1e-6, (name, np.median(relative))
    # learning rate 0 leaves the master parameters untouched
    np.testing.assert_array_equal(net.params.numpy(), params)


def test_train_step_is_deterministic_and_applies_adam(gsn, context):
    n = 20000
    features, labels, weights, params = _problem(n, 9)
    x, y = gsn.DeviceBuffer.from_numpy(features), gsn.DeviceBuffer.from_numpy(labels)
    runs = []
    for _ in range(2):
        net = gsn.ProposalNetwork(context, params)
        gradients = gsn.DeviceBuffer((mlp.PARAMETERS,), np.float32)
        net.train_step(x, y, n, 1e-3, gradients=gradients)
        context.synchronize()
        runs.append((gradients.numpy(), net.params.numpy(), net.first.numpy(), net.second.numpy(),
                     net.weights.numpy()))
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)
    g, new_params, first, second, weights_bf16 = runs[0]
    expected_params, expected_first, expected_second = mlp.adam_f32(params, np.zeros_like(params),
                                                                    np.zeros_like(params), g, 1e-3, 1)
    np.testing.assert_allclose(first, expected_first, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(second, expected_second, rtol=1e-5, atol=1e-18)
    np.testing.assert_allclose(new_params, expected_params, rtol=1e-6, atol=2e-7)
    np.testing.assert_array_equal(weights_bf16, mlp.bf16_weight_block(new_params))


def test_training_learns_a_threshold_rule(gsn, context):
    # Labels from a rule of the kind density control applies: clone when feature 0 >= 1 and
    # feature 5 <= 0, split when feature 0 >= 1 and feature 5 > 0, keep otherwise.
    rng = np.random.default_rng(4)
    n = 16384
    features = rng.standard_normal((n, mlp.IN)).astype(np.float32)
    features[:, 0] = rng.uniform(0.0, 2.0, n)
    labels = np.where(features[:, 0] >= 1.0, np.where(features[:, 5] > 0.0, 2, 1), 0).astype(np.int32)
    net = gsn.ProposalNetwork(context, mlp.init_params(4))
    x, y = gsn.DeviceBuffer.from_numpy(features), gsn.DeviceBuffer.from_numpy(labels)
    first_loss = net.train_step(x, y, n, 1e-2, want_loss=True)
    for _ in range(400):
        net.train_step(x, y, n, 1e-2)
    last_loss = net.train_step(x, y, n, 1e-2, want_loss=True)
    logits = gsn.DeviceBuffer((n, 3), np.float32)
    net.forward(x, n, logits)
    context.synchronize()
    accuracy = np.mean(np.argmax(logits.numpy(), axis=1) == labels)
    assert last_loss < 0.2 * first_loss
    assert accuracy > 0.97, accuracy
