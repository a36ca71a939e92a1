"""Stored-state rasteriser variant (final transmittance + last contributor per pixel, back-to-front
backward) against the recompute form and the FP64 oracle.

- The stored forward writes the same image, bit for bit, as the plain forward; its image matches
  the FP32 forward mirror, and its per-pixel state is in range (the backward below checks that
  the state replays the forward exactly: a wrong T or last index would break the gradients).
- Its screen-space gradients match the FP64 replay under the rule of test_gpu_solver.py
  (1e-5 of the error-bound magnitude on Gaussians no pixel marks ambiguous): recovering T by
  division has the same 1 / (1 - alpha) conditioning the magnitude already carries.
- Its parameter gradients agree with the recompute form within 2e-5 of each row's scale.
"""
import numpy as np
import pytest

from helpers import GpuRun, config0_camera

pytestmark = pytest.mark.gpu

SCREEN_TOLERANCE = 1e-5


def _scene(oracle, scenes, gsn, seed, count, width, sh_degree):
    params = scenes.config0_params(seed=seed, count=count, sh_degree=3)
    params[3:6] += np.float32(1.0)
    cam = config0_camera(gsn, width, width)
    truth = scenes.perturb(params, seed + 1, 0.01)
    image = oracle.render_reference(truth, cam, (0.1, 0.2, 0.3), sh_degree)
    target = np.clip(np.rint(image * 255.0), 0, 255).astype(np.uint8)
    return params, cam, target


def _stored_run(gsn, context, run, target, background):
    cam = run.camera
    pixels = cam.width * cam.height
    image = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    final_t = gsn.DeviceBuffer((pixels,), np.float32)
    last = gsn.DeviceBuffer((pixels,), np.int32)
    gsn.render_stored(context, run.projection, image, background, final_t, last)
    image_grad = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    gsn.image_gradients(context, cam, image, gsn.DeviceBuffer.from_numpy(target), image_grad)
    n = run.count
    gradients = gsn.DeviceBuffer.zeros((58, n), np.float32)
    screen = gsn.DeviceBuffer.zeros((11, n), np.float32)
    stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    abs_stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    gsn.gaussian_gradients_stored(context, run.projection, run.params, run.sh_degree, final_t, last, background,
                                  image_grad, gradients, stats, abs_stats, screen)
    context.synchronize()
    return {"image": image.numpy(), "image_grad": image_grad.numpy(),
            "final_t": final_t.numpy().reshape(cam.height, cam.width),
            "last": last.numpy().reshape(cam.height, cam.width), "gradients": gradients.numpy(),
            "screen": screen.numpy(), "statistics": stats.numpy()}


@pytest.mark.parametrize("sh_degree", [0, 3])
def test_stored_state_matches_recompute_and_fp64(gsn, context, oracle, scenes, sh_degree):
    background = (0.1, 0.2, 0.3)
    params, cam, target = _scene(oracle, scenes, gsn, 61 + sh_degree, 2000, 128, sh_degree)
    run = GpuRun(gsn, context, params, cam, background, sh_degree)
    stored = _stored_run(gsn, context, run, target, background)
    np.testing.assert_array_equal(stored["image"], run.image_host())

    # per-pixel state against the FP32 forward mirror
    f32 = oracle.project_f32(params, cam, sh_degree)
    lists = oracle.tile_lists(cam, f32)
    mirror, _ = oracle.render_f32(cam, f32, lists, background)
    np.testing.assert_allclose(stored["image"], mirror, rtol=1e-5, atol=3e-5)
    assert stored["last"].min() >= 0 and stored["final_t"].max() <= 1.0

    recompute = run.gradients(target)
    for row in range(58):
        scale = np.abs(recompute["gradients"][row]).max() + 1e-30
        assert np.abs(stored["gradients"][row] - recompute["gradients"][row]).max() <= 2e-5 * scale, row
    np.testing.assert_array_equal(stored["statistics"][1], recompute["statistics"][1])

    ref = oracle.raster_gradients_f64(cam, f32, background, target, image_grad=stored["image_grad"])
    keep = ~ref["tainted"]
    assert keep.mean() > 0.5
    for k in range(11):
        error = np.abs(stored["screen"][k] - ref["screen"][k])[keep]
        bound = SCREEN_TOLERANCE * ref["magnitude"][k][keep] + 1e-9 * np.abs(ref["screen"][k]).max()
        worst = np.argmax(error - bound)
        assert np.all(error <= bound), f"field {k}: error {error[worst]} bound {bound[worst]}"


def test_stored_forward_rejects_missing_state(gsn, context, scenes):
    params = scenes.config0_params(seed=3, count=100)
    cam = config0_camera(gsn, 64, 64)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), 0)
    with pytest.raises(gsn.GsnError):
        gsn.render_stored(context, run.projection, run.image, (0.0, 0.0, 0.0), None, None)
