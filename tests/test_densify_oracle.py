"""Pins of the learned-densification restatement (oracle/densify.py and the proposal form of the
density count in oracle.c) on hand-built cases whose answers follow from the definitions."""
import math

import numpy as np

from oracle import densify


def _params(n):
    params = np.zeros((58, n), np.float32)
    params[9] = 5.0  # opaque
    return params


def test_accumulate_radius_depth_and_max():
    n = 3
    state = np.zeros((6, n), np.float32)
    sigma = np.array([2.0, 0.5, 1.0])
    conics = np.zeros((n, 4), np.float32)
    conics[:, 0] = 1.0 / sigma ** 2
    conics[:, 2] = 1.0 / sigma ** 2
    conics[2, 2] = 1.0 / 16.0  # anisotropic: major sigma 4 along y
    stats = np.array([[1.0, 2.0, 3.0], [1.0, 0.0, 1.0]], np.float32)
    absolute = np.array([[4.0, 5.0, 6.0], [0, 0, 0]], np.float32)
    state = densify.accumulate(state, stats, absolute, np.array([2.0, 3.0, 4.0], np.float32), conics)
    state = densify.accumulate(state, stats * 0.5, absolute * 2, np.array([1.0, 3.0, 8.0], np.float32), conics)
    np.testing.assert_allclose(state[5], [6.0, 0.0, 12.0], rtol=1e-6)  # 3 sigma of the major axis
    np.testing.assert_array_equal(state[0], [1.0, 0.0, 3.0])             # max over views, unobserved untouched
    np.testing.assert_array_equal(state[1], [8.0, 0.0, 12.0])
    np.testing.assert_array_equal(state[2], [3.0, 0.0, 12.0])            # depth sum
    np.testing.assert_array_equal(state[3], [1.0, 0.0, 4.0])             # depth min
    np.testing.assert_array_equal(state[4], [2.0, 0.0, 8.0])             # depth max


def test_features_follow_the_rule_quantities():
    n = 2
    params = _params(n)
    params[3:6, 0] = [math.log(0.02), math.log(0.5), math.log(0.1)]
    params[9, 0] = 0.0
    params[10, 0], params[26, 0], params[42, 0] = 0.1, 0.2, 0.3
    sums = np.array([6e-4, 0.0], np.float32)
    absolute = np.array([9e-4, 0.0], np.float32)
    counts = np.array([3, 0], np.uint32)
    state = np.zeros((6, n), np.float32)
    state[:, 0] = [4e-4, 6e-4, 12.0, 2.0, 6.0, 1.0]
    f = densify.features(params, sums, absolute, counts, state, 2e-4, 0.1, 2.0)
    assert f.shape == (2, 16)
    np.testing.assert_allclose(f[0, 0], 1.0)                       # mean gradient at the threshold
    np.testing.assert_allclose(f[0, 1], 1.5)
    np.testing.assert_allclose(f[0, 2], math.log(3.0))
    np.testing.assert_allclose(f[0, 3], math.log(4.0))
    np.testing.assert_allclose(f[0, 4], math.log(4.0))
    np.testing.assert_allclose(f[0, 5:8], [math.log(5.0), 0.0, math.log(0.2)], atol=1e-6)  # descending
    np.testing.assert_allclose(f[0, 8], 0.5)
    np.testing.assert_allclose(f[0, 9:12], [math.log(2.0), 0.0, math.log(3.0)], atol=1e-7)  # mean 4, min 2, max 6 over 2
    np.testing.assert_allclose(f[0, 12:15], [0.1, 0.2, 0.3], rtol=1e-6)
    np.testing.assert_allclose(f[0, 15], math.log(2.0))
    # never observed: observation-derived entries vanish
    assert np.all(f[1, [0, 1, 2, 3, 4, 9, 10, 11, 15]] == 0.0)


def test_outcome_labels_hand_built():
    # 5 Gaussians before the step: 0 kept, 1 cloned, 2 split, 3 pruned, 4 cloned.
    # Outputs: [keep: 0, 1, 4 | clones: 1, 4 | split children 0: 2 | split children 1: 2]
    decisions = np.array([[1, 0, 0, 0], [1, 1, 0, 0], [0, 0, 1, 1], [0, 0, 0, 0], [1, 1, 0, 0], [0, 0, 0, 0]],
                         np.int32)
    offsets = np.zeros_like(decisions)
    offsets[1:] = np.cumsum(decisions[:-1], axis=0)
    assert list(offsets[5]) == [3, 2, 1, 1]
    m = 3 + 2 + 2
    params = np.zeros((58, m), np.float32)
    params[9] = 1.0
    counts = np.ones(m, np.uint32)
    logit = math.log(0.02 / 0.98)
    # clone of Gaussian 4 (output 4) became transparent; split child 1 of Gaussian 2 (output 6) unseen
    params[9, 4] = -6.0
    counts[6] = 0
    labels = densify.outcome_labels(decisions, offsets, params, counts, logit)
    np.testing.assert_array_equal(labels, [0, 1, 0, -1, 0])
    counts[6] = 1
    labels = densify.outcome_labels(decisions, offsets, params, counts, logit)
    np.testing.assert_array_equal(labels, [0, 1, 2, -1, 0])


def test_proposed_count_takes_argmax_and_keeps_pruning(oracle):
    n = 6
    params = _params(n)
    params[3:6] = math.log(0.05)
    params[3, 4] = math.log(0.5)     # large: oversize-pruned when split
    params[9, 5] = -8.0              # transparent: pruned whatever the proposal
    logits = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 1, 1], [0, 0, 1], [0, 5, 0]], np.float32)
    decisions, offsets = oracle.density_count_proposed_f32(params, logits, 0.1, 0.005, 0.2, True)
    np.testing.assert_array_equal(decisions[:n], [[1, 0, 0, 0], [1, 1, 0, 0], [0, 0, 1, 1], [1, 1, 0, 0],
                                                  [0, 0, 1, 0], [0, 0, 0, 0]])
    np.testing.assert_array_equal(offsets[n], decisions[:n].sum(axis=0))
    # ties and NaN favour keep
    logits = np.array([[1, 1, 1], [np.nan, 2, 3], [0, np.nan, 0]], np.float32)
    decisions, _ = oracle.density_count_proposed_f32(params[:, :3].copy(), logits, 0.1, 0.005, 0.2, False)
    np.testing.assert_array_equal(decisions[:3, :2], [[1, 0], [1, 0], [1, 0]])
