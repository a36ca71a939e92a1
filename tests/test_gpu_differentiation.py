"""Differentiation check at the config-2 size: 2,000 Gaussians with the config-0 distributions at
SH degree 3, rendered at 512x512 from 8 random cameras (fov 60 degrees, distance 3). The GPU's
hand-written backward (summed over the 8 views) is compared

  1. with the oracle's FP64 analytic gradient (the CPU reference gradient), on Gaussians that no
     view marks ambiguous: |gpu - fp64| <= 2e-4 |fp64| + 1e-4 max|row| per parameter row (the
     device projects in FP32: pixel-space means carry ~|m| * 6e-8 px of absolute error, which
     perturbs each pixel's term by ~1e-5 relative, and a parameter row sums terms of both signs);
  2. with FP64 central finite differences (h = 1e-4) of the summed L1+SSIM loss, for the means,
     log-scales, axis-angle rotations, opacity logits and SH coefficients (DC and one degree-3
     term): |gpu - fd| <= 2e-3 |fd| + 1e-4 max|row|.
The rendered loss is only piecewise smooth: a pixel's alpha jumps from 0 to 1/255 at the cutoff,
tile rectangles are discrete and depth order can swap, so a difference quotient whose window
straddles such a jump measures the jump, not the derivative (with 2,000 overlapping splats that
happens for a good share of the geometry parameters at h = 1e-4). A pick is therefore compared only
if its quotient is stable, |fd(h) - fd(h/2)| <= 1e-3 |fd(h)| + 1e-5 max|row|; every parameter
group must keep at least one stable pick. Positions use h = 1e-6 instead: at 512 px and distance 3
a unit of position is ~150 px, so h = 1e-4 moves every splat edge by 0.015 px and crosses the
cutoff and L1 kinks (|I - T| changes sign) in a large share of the covered pixels; at 1e-6 the
quotients are stable to six digits. The FP64 loss resolves such steps (relative 1e-16).
"""
import concurrent.futures
import math

import numpy as np
import pytest

from helpers import GpuRun

pytestmark = pytest.mark.gpu

VIEWS = 8
SIZE = 512


def _cameras(gsn, scenes, seed):
    focal = scenes.fov_focal(SIZE, 60.0)
    return [gsn.Camera.make(SIZE, SIZE, focal, focal, 0.5 * (SIZE - 1), 0.5 * (SIZE - 1), R, t)
            for R, t in scenes.sphere_cameras(VIEWS, 3.0, seed)]


def _loss_sum(oracle, params, cams, targets):
    total = 0.0
    for cam, target in zip(cams, targets):
        total += oracle.loss_f64(oracle.render_reference(params, cam, (0.0, 0.0, 0.0), 3), target)
    return total


def test_backward_matches_cpu_gradient_and_finite_differences(gsn, context, oracle, scenes):
    params = scenes.config0_params(seed=21, count=2000, sh_degree=3)
    truth = scenes.perturb(params, 22, 0.01)
    cams = _cameras(gsn, scenes, 23)
    targets = [np.clip(np.rint(oracle.render_reference(truth, cam, (0.0, 0.0, 0.0), 3) * 255.0), 0, 255)
               .astype(np.uint8) for cam in cams]
    gpu = np.zeros((58, params.shape[1]))
    cpu = np.zeros((58, params.shape[1]))
    tainted = np.zeros(params.shape[1], bool)
    for cam, target in zip(cams, targets):
        run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), 3)
        gpu += run.gradients(target, screen=False)["gradients"]
        f32 = oracle.project_f32(params, cam, 3)
        ref = oracle.gradients_f64(params, cam, 3, (0.0, 0.0, 0.0), target, rects_from=f32)
        cpu += ref["gradients"]
        tainted |= ref["tainted"]
    keep = ~tainted
    assert keep.mean() > 0.3
    for row in range(58):
        error = np.abs(gpu[row] - cpu[row])[keep]
        bound = 2e-4 * np.abs(cpu[row])[keep] + 1e-4 * np.abs(cpu[row]).max() + 1e-30
        assert np.all(error <= bound), f"row {row}: worst {np.max(error - bound)}"

    # finite differences of the summed FP64 loss
    groups = {"position": [0, 1, 2], "log-scale": [3, 4, 5], "rotation": [6, 7, 8], "opacity": [9],
              "sh": [10, 26, 42, 25]}
    steps = {"position": 1e-6}

    def central(task):
        row, g, step_size = task
        high, low = params.copy(), params.copy()
        high[row, g] += np.float32(step_size)
        low[row, g] -= np.float32(step_size)
        step = float(high[row, g]) - float(low[row, g])
        return (_loss_sum(oracle, high, cams, targets) - _loss_sum(oracle, low, cams, targets)) / step

    compared = {}
    with concurrent.futures.ThreadPoolExecutor(max_workers=16) as pool:
        for group, rows in groups.items():
            compared[group] = 0
            for row in rows:
                candidates = np.flatnonzero(keep)
                best = [int(g) for g in candidates[np.argsort(-np.abs(cpu[row][candidates]))[:6]]]
                h = steps.get(group, 1e-4)
                tasks = [(row, g, step) for g in best for step in (h, 0.5 * h)]
                values = list(pool.map(central, tasks))
                scale = np.abs(gpu[row]).max()
                for k, g in enumerate(best):
                    fd, fd_half = values[2 * k], values[2 * k + 1]
                    assert math.isfinite(fd) and math.isfinite(fd_half)
                    if abs(fd - fd_half) > 1e-3 * abs(fd) + 1e-5 * scale:
                        continue  # the window straddles a discontinuity of the rendered loss
                    assert abs(gpu[row, g] - fd) <= 2e-3 * abs(fd) + 1e-4 * scale, (row, g, gpu[row, g], fd)
                    compared[group] += 1
    print(f"finite-difference picks compared per group: {compared}", flush=True)
    assert all(count >= 1 for count in compared.values()), compared
