"""Synthetic training workloads shared by bench.py and the GPU tests.

build_workload: a ground-truth scene (the 64-blob mixture, or the shell + ground plane), its
log-scales from the GPU 3-NN initialisation, cameras on a sphere or hemisphere looking at the
origin, u8 targets rendered from the ground truth on the GPU and kept resident, and the trained
state: the ground truth perturbed by N(0, 0.01) (seeded).
"""
from __future__ import annotations

import sys

import numpy as np

if __package__:
    from . import scenes
else:  # loaded by path
    scenes = sys.modules["gs_native_amd.scenes"]


def build_workload(gsn, context, gaussians: int, width: int, height: int, cameras: int, seed: int = 0,
                   scene: str = "blobs", camera_radius: float = 5.0, hemisphere: bool = False):
    """Returns (params, cameras, targets, scene extent, visible Gaussians per camera)."""
    if scene == "shell":
        truth = scenes.shell_plane_scene(gaussians, seed=seed, sh_degree=3)
    else:
        truth = scenes.blob_scene(gaussians, seed=seed, sh_degree=3)
    _, mean_dist2 = gsn.knn3(context, truth[0:3].T.copy())
    truth[3:6] = scenes.scales_from_mean_dist2(mean_dist2)[None, :]
    params = scenes.perturb(truth, seed + 1, 0.01)
    focal = scenes.fov_focal(width, 60.0)
    poses = scenes.sphere_cameras(cameras, camera_radius, seed + 2, hemisphere=hemisphere)
    cams = [gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t) for R, t in poses]
    centres = np.stack([-R.T @ t for R, t in poses])
    extent = 1.1 * float(np.linalg.norm(centres - centres.mean(0), axis=1).max())
    device_truth = gsn.DeviceBuffer.from_numpy(truth)
    image = gsn.DeviceBuffer((height, width, 3), np.float32)
    targets, visible = [], []
    projection = None
    for cam in cams:
        projection = gsn.render(context, device_truth, gaussians, cam, (0.0, 0.0, 0.0), image, 3, projection)
        target = np.clip(np.rint(image.numpy() * 255.0), 0, 255).astype(np.uint8)
        targets.append(gsn.DeviceBuffer.from_numpy(target))
    del device_truth
    device_params = gsn.DeviceBuffer.from_numpy(params)
    for cam in cams:
        projection = gsn.render(context, device_params, gaussians, cam, (0.0, 0.0, 0.0), image, 3, projection)
        counts = np.diff(projection.record_offsets.numpy())
        visible.append(int((counts > 0).sum()))
    del device_params
    return params, cams, targets, extent, visible


def initial_from_points(gsn, context, points: np.ndarray) -> np.ndarray:
    """The reference initialisation from a point cloud (gaussian_splatting.cpp:95-128): positions
    of `points` (a [58][n] scene of which rows 0-2 and the SH DC rows are used), log-scales from the
    mean 3-NN squared distance (GPU LBVH), identity rotation, opacity 0.1, SH DC from the point
    colour, higher bands zero."""
    count = points.shape[1]
    init = np.zeros((58, count), np.float32)
    init[0:3] = points[0:3]
    _, d2 = gsn.knn3(context, points[0:3].T.copy())
    init[3:6] = scenes.scales_from_mean_dist2(d2)[None, :]
    init[9] = np.float32(np.log(0.1 / 0.9))
    for c in range(3):
        init[10 + 16 * c] = points[10 + 16 * c]
    return init
