"""The backward stamps every record it reaches with a per-step epoch byte (1..255) that the
per-Gaussian pass compares against; the flag buffer is cleared when it is allocated and when the
epoch wraps. These tests drive one context through more than 255 gradient steps, alternating
two projections with different record counts, and require every step to reproduce that
projection's first result bit for bit: a stale stamp that matched the current epoch would add a
stale partial, and a missing stamp would drop one.
"""
import numpy as np
import pytest

from helpers import config0_camera

pytestmark = pytest.mark.gpu

STEPS = 300  # past one epoch wrap (255 steps)


def _projection(gsn, context, params, cam, sh_degree):
    device_params = gsn.DeviceBuffer.from_numpy(params)
    image = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    projection = gsn.render(context, device_params, params.shape[1], cam, (0.0, 0.0, 0.0), image, sh_degree)
    return device_params, image, projection


def _gradients(gsn, context, run, image_grad, sh_degree):
    device_params, image, projection = run
    n = projection.count
    gradients = gsn.DeviceBuffer.zeros((58, n), np.float32)
    stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    abs_stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    gsn.gaussian_gradients(context, projection, device_params, sh_degree, image, image_grad, gradients, stats,
                           abs_stats)
    context.synchronize()
    return gradients.numpy(), stats.numpy()


def test_epoch_wrap_and_alternating_projections(gsn, scenes, oracle):
    context = gsn.Context()  # fresh epoch counter and flag buffer
    sh_degree = 3
    runs, image_grads = [], []
    for seed, count, width in ((71, 3000, 128), (72, 1500, 96)):
        params = scenes.config0_params(seed=seed, count=count, sh_degree=3)
        params[3:6] += np.float32(1.0)
        cam = config0_camera(gsn, width, width)
        truth = scenes.perturb(params, seed + 1, 0.01)
        target = np.clip(np.rint(oracle.render_reference(truth, cam, (0, 0, 0), sh_degree) * 255.0), 0,
                         255).astype(np.uint8)
        run = _projection(gsn, context, params, cam, sh_degree)
        image_grad = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
        gsn.image_gradients(context, cam, run[1], gsn.DeviceBuffer.from_numpy(target), image_grad)
        runs.append(run)
        image_grads.append(image_grad)
    assert runs[0][2].record_count != runs[1][2].record_count
    expected = [_gradients(gsn, context, runs[k], image_grads[k], sh_degree) for k in range(2)]
    for k in range(2):
        assert np.abs(expected[k][0]).max() > 0
    for step in range(STEPS):
        k = step % 2
        gradients, stats = _gradients(gsn, context, runs[k], image_grads[k], sh_degree)
        np.testing.assert_array_equal(gradients.view(np.uint32), expected[k][0].view(np.uint32), err_msg=f"step {step}")
        np.testing.assert_array_equal(stats.view(np.uint32), expected[k][1].view(np.uint32), err_msg=f"step {step}")
