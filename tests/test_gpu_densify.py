"""Learned densification inputs and policy on the GPU against oracle/densify.py and the C oracle:
feature state and features (tolerance: float32 transcendentals, 2e-6 relative + 2e-6 absolute on
O(1) features), outcome labels and proposal-driven density decisions (bit-exact), and the trainer
integration (sample logging and a learned policy run end to end)."""
import ctypes
import math

import numpy as np
import pytest

from oracle import densify

pytestmark = pytest.mark.gpu


def _random_params(rng, n):
    params = np.zeros((58, n), np.float32)
    params[0:3] = rng.uniform(-1, 1, (3, n))
    params[3:6] = rng.normal(-4.0, 1.0, (3, n))
    params[6:9] = rng.normal(0, 1, (3, n))
    params[9] = rng.normal(0, 3, n)
    params[10:58] = rng.normal(0, 0.3, (48, n))
    return params


def test_feature_state_and_features_match_oracle(gsn, context):
    rng = np.random.default_rng(5)
    n = 5000
    d = gsn.DeviceBuffer
    state = d((6, n), np.float32)
    gsn.call("gsn_density_features_clear", context.handle, state.ptr, n)
    expected_state = np.zeros((6, n), np.float32)
    for view in range(4):
        stats = np.zeros((2, n), np.float32)
        stats[0] = rng.exponential(3e-4, n)
        stats[1] = (rng.uniform(size=n) < 0.6).astype(np.float32)
        absolute = np.zeros((2, n), np.float32)
        absolute[0] = stats[0] * rng.uniform(1.0, 2.0, n)
        depths = rng.uniform(0.5, 8.0, n).astype(np.float32)
        conics = np.zeros((n, 4), np.float32)
        conics[:, 0] = rng.uniform(0.01, 2.0, n)
        conics[:, 2] = rng.uniform(0.01, 2.0, n)
        conics[:, 1] = rng.uniform(-0.9, 0.9, n) * np.sqrt(conics[:, 0] * conics[:, 2])
        conics[:, 3] = rng.uniform(0.01, 1, n)
        buffers = [d.from_numpy(a) for a in (stats, absolute, depths, conics)]
        gsn.call("gsn_density_features_accumulate", context.handle, *[b.ptr for b in buffers], n, state.ptr, n)
        expected_state = densify.accumulate(expected_state, stats, absolute, depths, conics)
    context.synchronize()
    actual_state = state.numpy()
    exact_rows = [0, 1, 3, 4]  # max/min selections are exact
    np.testing.assert_array_equal(actual_state[exact_rows], expected_state[exact_rows])
    np.testing.assert_allclose(actual_state[2], expected_state[2], rtol=1e-6)
    # the radius takes the smaller eigenvalue of the conic, half trace - sqrt(...), which cancels
    np.testing.assert_allclose(actual_state[5], expected_state[5], rtol=1e-4)

    params = _random_params(rng, n)
    counts = rng.integers(0, 5, n).astype(np.uint32)
    sums = (rng.exponential(3e-4, n) * counts).astype(np.float32)
    absolute_sums = (sums * rng.uniform(1, 2, n)).astype(np.float32)
    features = d((n, 16), np.float32)
    inputs = [d.from_numpy(a) for a in (params, sums, absolute_sums, counts)]
    gsn.call("gsn_density_features", context.handle, inputs[0].ptr, n, inputs[1].ptr, inputs[2].ptr, inputs[3].ptr,
             state.ptr, n, 2e-4, 0.02, 3.0, features.ptr)
    context.synchronize()
    expected = densify.features(params, sums, absolute_sums, counts, actual_state, 2e-4, 0.02, 3.0)
    np.testing.assert_allclose(features.numpy(), expected, rtol=2e-6, atol=2e-6)


def test_proposed_density_count_bit_exact(gsn, context, oracle):
    rng = np.random.default_rng(6)
    n = 20000
    params = _random_params(rng, n)
    params[3:6] = rng.normal(-3.0, 1.2, (3, n))
    logits = rng.normal(0, 1, (n, 3)).astype(np.float32)
    logits[::17, 1] = logits[::17, 0]  # ties
    logits[::23, 2] = np.nan
    d = gsn.DeviceBuffer
    zeros = d.zeros((n,), np.float32)
    counts = d.zeros((n,), np.uint32)
    decisions = d((n + 1, 4), np.int32)
    offsets = d((n + 1, 4), np.int32)
    totals = (ctypes.c_int32 * 4)()
    p, l = d.from_numpy(params), d.from_numpy(logits)
    for prune_large in (0, 1):
        gsn.call("gsn_density_count_proposed", context.handle, p.ptr, n, zeros.ptr, zeros.ptr, counts.ptr, l.ptr,
                 0.05, 0.005, 0.3, prune_large, decisions.ptr, offsets.ptr, totals)
        expected_d, expected_o = oracle.density_count_proposed_f32(params, logits, 0.05, 0.005, 0.3, prune_large)
        np.testing.assert_array_equal(decisions.numpy(), expected_d)
        np.testing.assert_array_equal(offsets.numpy(), expected_o)
        assert list(totals) == list(expected_o[n])


def test_outcome_labels_bit_exact(gsn, context, oracle):
    rng = np.random.default_rng(7)
    n = 8000
    params = _random_params(rng, n)
    sums = rng.exponential(2e-4, n).astype(np.float32)
    counts = rng.integers(0, 3, n).astype(np.uint32)
    decisions, offsets = oracle.density_count_f32(params, sums * counts, sums, counts, 2e-4, 0.02, 0.005, 0.3, False,
                                                  0.0)
    t = offsets[n]
    m = int(t[0] + t[1] + 2 * t[3])
    after = _random_params(rng, m)
    counts_after = rng.integers(0, 2, m).astype(np.uint32)
    logit = math.log(0.02 / 0.98)
    # keep opacities away from the threshold so that host and device agree on its float32 value
    close = np.abs(after[9] - logit) < 1e-3
    after[9, close] += 0.01
    d = gsn.DeviceBuffer
    labels = d((n,), np.int32)
    # every device buffer stays referenced until the call has completed
    inputs = [d.from_numpy(a) for a in (decisions, offsets, after, counts_after)]
    gsn.call("gsn_density_outcomes", context.handle, inputs[0].ptr, inputs[1].ptr,
             (ctypes.c_int32 * 4)(*[int(v) for v in t]), n, inputs[2].ptr, inputs[3].ptr, m, 0.02, labels.ptr)
    context.synchronize()
    expected = densify.outcome_labels(decisions, offsets, after, counts_after, logit)
    np.testing.assert_array_equal(labels.numpy(), expected)
    assert set(np.unique(expected)) >= {-1, 0, 1, 2}


def _small_scene(gsn, scenes, context, count, size, cameras):
    truth = scenes.blob_scene(count, seed=3, sh_degree=3)
    _, mean_dist2 = gsn.knn3(context, truth[0:3].T.copy())
    truth[3:6] = scenes.scales_from_mean_dist2(mean_dist2)[None, :]
    params = scenes.perturb(truth, 4, 0.02)
    focal = scenes.fov_focal(size, 60.0)
    poses = scenes.sphere_cameras(cameras, 5.0, 5)
    cams = [gsn.Camera.make(size, size, focal, focal, 0.5 * (size - 1), 0.5 * (size - 1), R, t) for R, t in poses]
    image = gsn.DeviceBuffer((size, size, 3), np.float32)
    device_truth = gsn.DeviceBuffer.from_numpy(truth)
    targets = []
    projection = None
    for cam in cams:
        projection = gsn.render(context, device_truth, count, cam, (0.0, 0.0, 0.0), image, 3, projection)
        targets.append(np.clip(np.rint(image.numpy() * 255.0), 0, 255).astype(np.uint8))
    return params, cams, targets


def test_trainer_logs_outcomes_and_runs_a_policy(gsn, context, scenes):
    import gs_native_amd.trainer as trainer_module
    from oracle import mlp
    params, cams, targets = _small_scene(gsn, scenes, context, 20000, 128, 8)
    log = trainer_module.OutcomeLog()
    trainer = trainer_module.Trainer(context, params, cams, targets, 5.0, outcome_log=log, useful_opacity=(0.02, 0.2))
    for _ in range(800):
        trainer.iteration()
    context.synchronize()
    features, labels, weights = log.arrays()
    assert features.shape[1] == 16 and features.shape[0] == labels.shape[0] > 0
    assert np.all(np.isfinite(features))
    assert set(np.unique(labels)) <= {0, 1, 2}
    assert np.all(weights == 1.0)
    # a stricter usefulness threshold only turns densify labels into keep labels
    _, strict, _ = log.arrays(1)
    assert np.all((strict == 0) | (strict == labels))
    assert np.sum(strict > 0) <= np.sum(labels > 0)
    # a network that always answers keep: no Gaussian is added at any density step
    parts = {name: np.zeros(shape, np.float32) for name, shape in mlp.SHAPES}
    parts["b3"][:] = [1.0, 0.0, 0.0]
    network = gsn.ProposalNetwork(context, mlp.pack(parts))
    keep_all = trainer_module.Trainer(context, params, cams, targets, 5.0, policy=network)
    for _ in range(800):
        keep_all.iteration()
    context.synchronize()
    assert keep_all.history and all(changes[0] == 0 for _, _, _, changes in keep_all.history)
    assert keep_all.count <= params.shape[1]
    # a network that always answers split proposes far more growth than the rule: the growth guard
    # hands every density step back to the rule, so the run repeats the rule run exactly
    parts["b3"][:] = [0.0, 0.0, 1.0]
    split_all = trainer_module.Trainer(context, params, cams, targets, 5.0,
                                       policy=gsn.ProposalNetwork(context, mlp.pack(parts)))
    for _ in range(800):
        split_all.iteration()
    context.synchronize()
    assert split_all.policy_fallbacks == len(split_all.history) > 0
    assert split_all.history == trainer.history
    assert split_all.count == trainer.count


def test_policy_runs_on_the_trainers_stream(gsn, context, scenes):
    """A network made on one context drives trainers on contexts of their own: its forward pass
    runs on the trainer's stream (after the features, before the decisions read the logits), so
    two such runs agree bit for bit."""
    import gs_native_amd.trainer as trainer_module
    from oracle import mlp
    params, cams, targets = _small_scene(gsn, scenes, context, 20000, 128, 8)
    parts = mlp.unpack(mlp.init_params(3))
    parts["b3"][:] = [0.0, 0.2, -0.2]
    network = gsn.ProposalNetwork(context, mlp.pack(parts))
    runs = []
    for _ in range(2):
        own = gsn.Context()
        trainer = trainer_module.Trainer(own, params, cams, targets, 5.0, policy=network)
        for _ in range(800):
            trainer.iteration()
        own.synchronize()
        runs.append((trainer.history, trainer.params_host().copy()))
        del trainer
        own.close()
    assert runs[0][0] == runs[1][0]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
