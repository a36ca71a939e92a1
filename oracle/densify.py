"""NumPy restatement of the learned-densification inputs: the per-view feature state, the
16 features and the outcome labels.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP kernels. The product
never imports it. These components have no counterpart in the reference; the gradient means
reuse its density statistics (density_controller.cu:83, 97).
"""
from __future__ import annotations

import numpy as np

STATE_ROWS = 6
FEATURES = 16


def accumulate(state: np.ndarray, statistics: np.ndarray, absolute_statistics: np.ndarray, depths: np.ndarray,
               conics: np.ndarray) -> np.ndarray:
    """One view folded into state[6][N] (returns a new array)."""
    state = np.array(state, np.float32, copy=True)
    observed = statistics[1] != 0.0
    a, b, c = (conics[:, k].astype(np.float64) for k in range(3))
    smallest = 0.5 * (a + c) - np.sqrt((0.5 * (a - c)) ** 2 + b * b)
    with np.errstate(divide="ignore", invalid="ignore"):
        radius = np.where(smallest > 1e-8, 3.0 / np.sqrt(np.maximum(smallest, 1e-30)), 1e4)
    radius = np.minimum(radius, 1e4).astype(np.float32)
    d = depths.astype(np.float32)
    o = observed
    state[0, o] = np.maximum(state[0, o], statistics[0, o])
    state[1, o] = np.maximum(state[1, o], absolute_statistics[0, o])
    state[2, o] = state[2, o] + d[o]
    state[3, o] = np.where(state[3, o] > 0.0, np.minimum(state[3, o], d[o]), d[o])
    state[4, o] = np.maximum(state[4, o], d[o])
    state[5, o] = np.maximum(state[5, o], radius[o])
    return state


def features(params: np.ndarray, sums: np.ndarray, absolute_sums: np.ndarray, counts: np.ndarray,
             state: np.ndarray, gradient_threshold: float, split_threshold: float, extent: float) -> np.ndarray:
    """features[N][16] in float64 (compare with a tolerance)."""
    params = params.astype(np.float64)
    n = params.shape[1]
    counts = counts.astype(np.float64)
    observed = counts > 0
    inverse = np.where(observed, 1.0 / np.maximum(counts, 1.0), 0.0)
    f = np.zeros((n, FEATURES))
    f[:, 0] = sums * inverse / gradient_threshold
    f[:, 1] = absolute_sums * inverse / gradient_threshold
    f[:, 2] = np.where(observed, np.log1p(state[0] / gradient_threshold), 0.0)
    f[:, 3] = np.where(observed, np.log1p(state[1] / gradient_threshold), 0.0)
    f[:, 4] = np.log1p(counts)
    scales = np.sort(params[3:6], axis=0)[::-1]
    f[:, 5:8] = (scales - np.log(split_threshold)).T
    f[:, 8] = 1.0 / (1.0 + np.exp(-params[9]))
    with np.errstate(divide="ignore", invalid="ignore"):
        f[:, 9] = np.where(observed, np.log(state[2] * inverse / extent), 0.0)
        f[:, 10] = np.where(observed, np.log(state[3] / extent), 0.0)
        f[:, 11] = np.where(observed, np.log(state[4] / extent), 0.0)
    f[:, 12] = params[10]
    f[:, 13] = params[26]
    f[:, 14] = params[42]
    f[:, 15] = np.where(observed, np.log1p(state[5]), 0.0)
    return f


def outcome_labels(decisions: np.ndarray, offsets: np.ndarray, params_after: np.ndarray, counts_after: np.ndarray,
                   min_opacity_logit: float) -> np.ndarray:
    """Labels of the density step described by decisions/offsets ([N+1][4] int) given the
    Gaussians it produced after one interval: -1 pruned, 0 kept, 1/2 densified with every
    output still observed and at least the minimum opacity."""
    n = decisions.shape[0] - 1
    totals = offsets[n]
    m = params_after.shape[1]

    def useful(j):
        j = np.asarray(j)
        inside = (j >= 0) & (j < m)
        jj = np.clip(j, 0, max(m - 1, 0))
        return inside & (counts_after[jj] > 0) & (params_after[9, jj] >= np.float32(min_opacity_logit))

    d, o = decisions[:n], offsets[:n]
    labels = np.zeros(n, np.int32)
    removed = (d[:, 0] == 0) & (d[:, 1] == 0) & (d[:, 3] == 0)
    clone = d[:, 1] != 0
    split = (d[:, 3] != 0) & ~clone
    labels[removed] = -1
    clone_ok = useful(o[:, 0]) & useful(totals[0] + o[:, 1])
    labels[clone & clone_ok] = 1
    base = totals[0] + totals[1] + o[:, 3]
    split_ok = useful(base) & useful(base + totals[3])
    labels[split & split_ok & ~removed] = 2
    return labels
