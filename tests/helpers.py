"""Shared helpers of the parity tests: run the HIP path through the C ABI and the oracle on the
same seeded inputs."""
from __future__ import annotations

import numpy as np


def camera(gsn, width, height, fx, fy, cx, cy, rotation=None, translation=None, near=0.2, far=None):
    far = gsn.FLT_MAX if far is None else far
    return gsn.Camera.make(width, height, fx, fy, cx, cy, rotation, translation, near, far)


def reference_camera(gsn, width, height, rotation=None):
    """test_support.h:132-148 (MakeCamera): fx 20, fy 21, centred principal point."""
    return camera(gsn, width, height, 20.0, 21.0, 0.5 * (width - 1), 0.5 * (height - 1), rotation, None, 0.2, 100.0)


def config0_camera(gsn, width=256, height=256, fov=60.0, distance=3.0):
    f = 0.5 * width / np.tan(np.radians(fov) / 2.0)
    return camera(gsn, width, height, f, f, 0.5 * (width - 1), 0.5 * (height - 1), None, [0.0, 0.0, distance])


class GpuRun:
    """One forward (+ optional backward) of the HIP path on a parameter array."""

    def __init__(self, gsn, context, params: np.ndarray, cam, background=(0.0, 0.0, 0.0), sh_degree=0,
                 keep_keys=False):
        self.gsn = gsn
        self.context = context
        self.params_host = np.ascontiguousarray(params, np.float32)
        self.count = self.params_host.shape[1]
        self.camera = cam
        self.sh_degree = sh_degree
        self.background = background
        self.params = gsn.DeviceBuffer.from_numpy(self.params_host) if self.count else gsn.DeviceBuffer((1,), np.float32)
        self.image = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
        self.projection = gsn.render(context, self.params, self.count, cam, background, self.image, sh_degree,
                                     keep_keys=keep_keys)
        context.synchronize()

    def image_host(self) -> np.ndarray:
        return self.image.numpy()

    def lists(self) -> dict:
        p = self.projection
        r = p.record_count
        out = {
            "record_count": r,
            "record_offsets": p.record_offsets.numpy(),
            "tile_offsets": p.tile_offsets.numpy()[: self.camera.tile_count + 1],
            "tile_gaussian_ids": p.tile_gaussian_ids.numpy()[:r],
        }
        if p.sorted_keys is not None:
            out["sorted_keys"] = p.sorted_keys.numpy()[:r]
        return out

    def projected(self) -> dict:
        p = self.projection
        return {"means": p.means.numpy()[: self.count], "conic_opacities": p.conic_opacities.numpy()[: self.count],
                "colors": p.colors.numpy()[: self.count], "depths": p.depths.numpy()[: self.count],
                "tile_rects": p.tile_rects.numpy()[: self.count]}

    def gradients(self, target: np.ndarray, screen=True) -> dict:
        gsn = self.gsn
        n = max(self.count, 1)
        target_dev = gsn.DeviceBuffer.from_numpy(np.ascontiguousarray(target, np.uint8))
        image_grad = gsn.DeviceBuffer((self.camera.height, self.camera.width, 3), np.float32)
        gsn.image_gradients(self.context, self.camera, self.image, target_dev, image_grad)
        gradients = gsn.DeviceBuffer.zeros((58, n), np.float32)
        stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
        abs_stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
        screen_dev = gsn.DeviceBuffer.zeros((11, n), np.float32) if screen else None
        gsn.gaussian_gradients(self.context, self.projection, self.params, self.sh_degree, self.image, image_grad,
                               gradients, stats, abs_stats, screen_dev)
        self.context.synchronize()
        out = {"image_grad": image_grad.numpy(), "gradients": gradients.numpy()[:, : self.count],
               "statistics": stats.numpy()[:, : self.count], "absolute_statistics": abs_stats.numpy()[:, : self.count]}
        if screen:
            out["screen"] = screen_dev.numpy()[:, : self.count]
        return out


def solver_step(gsn, context, params: np.ndarray, cam, target: np.ndarray, sh_degree: int, background=(0.0, 0.0, 0.0),
                position_lr=1e-4, first=None, second=None, step=1):
    """Render + Solver::Iterate (fused gradients + Adam); returns updated params and moments."""
    n = params.shape[1]
    params_dev = gsn.DeviceBuffer.from_numpy(params)
    image = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    projection = gsn.render(context, params_dev, n, cam, background, image, sh_degree)
    target_dev = gsn.DeviceBuffer.from_numpy(np.ascontiguousarray(target, np.uint8))
    image_grad = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    gsn.image_gradients(context, cam, image, target_dev, image_grad)
    zeros = np.zeros((58, n), np.float32)
    moments = gsn.Moments.from_numpy(zeros if first is None else first, zeros if second is None else second)
    stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    abs_stats = gsn.DeviceBuffer.zeros((2, n), np.float32)
    c1, c2 = gsn.adam_corrections(step)
    gsn.gaussian_gradients_adam(context, projection, params_dev, sh_degree, image, image_grad, moments,
                                position_lr, c1, c2, stats, abs_stats)
    context.synchronize()
    first_out, second_out = moments.numpy()
    return {"params": params_dev.numpy(), "first": first_out, "second": second_out,
            "statistics": stats.numpy(), "absolute_statistics": abs_stats.numpy(), "image": image.numpy()}


def relative_error(actual: np.ndarray, expected: np.ndarray, scale: np.ndarray) -> np.ndarray:
    return np.abs(actual - expected) / np.maximum(scale, 1e-30)
