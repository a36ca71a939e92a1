"""Seeded synthetic scenes and cameras.

make_gaussians / make_camera / make_scene restate the reference's test fixtures
(test/test_support.h:113-162); config0 / blob_scene generate the benchmark workloads
(10k Gaussians at 256x256; the blob-mixture scenes used at 3M Gaussians and 1920x1080).
Parameters use the reference SoA layout float[58][N].
"""
from __future__ import annotations

import math

import numpy as np

SH_C0 = 0.28209479177387814
PARAM_COUNT = 58


def empty_params(count: int) -> np.ndarray:
    return np.zeros((PARAM_COUNT, count), np.float32)


def make_gaussians(count: int) -> np.ndarray:
    """test_support.h:113-130 (MakeGaussians)."""
    p = empty_params(count)
    i = np.arange(count)
    p[0] = ((i % 16).astype(np.float32) - np.float32(7.5)) * np.float32(0.035)
    p[1] = (((i // 16) % 16).astype(np.float32) - np.float32(7.5)) * np.float32(0.035)
    p[2] = np.float32(3.0) + np.float32(0.001) * i.astype(np.float32)
    p[3] = np.log(np.float32(0.12))
    p[4] = np.log(np.float32(0.16))
    p[5] = np.log(np.float32(0.2))
    p[6], p[7], p[8] = np.float32(0.15), np.float32(-0.2), np.float32(0.1)
    p[9] = np.float32(-1.0)
    for c in range(3):
        p[10 + 16 * c] = (np.float32(0.6) - np.float32(0.1) * np.float32(c) - np.float32(0.5)) / np.float32(SH_C0)
        for k in range(1, 16):
            p[10 + 16 * c + k] = np.float32(0.02) * np.sin((k + c + i).astype(np.float32))
    return p


def make_camera_args(width: int, height: int) -> dict:
    """test_support.h:132-148 (MakeCamera)."""
    return {"width": width, "height": height, "fx": 20.0, "fy": 21.0, "cx": 0.5 * (width - 1),
            "cy": 0.5 * (height - 1), "near": 0.2, "far": 100.0}


def axis_angle_matrix(axis: np.ndarray, angle: float) -> np.ndarray:
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * k + (1 - math.cos(angle)) * k @ k


def differentiable_scene_rotation() -> np.ndarray:
    """solver_test.cu:147-158: R = Ry(0.55) Rx(0.35) Rz(0.2)."""
    return (axis_angle_matrix([0, 1, 0], 0.55) @ axis_angle_matrix([1, 0, 0], 0.35)
            @ axis_angle_matrix([0, 0, 1], 0.2)).astype(np.float32)


def random_rotations(rng: np.random.Generator, count: int) -> np.ndarray:
    axis = rng.normal(size=(count, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    angle = rng.uniform(0.0, math.pi, size=count)
    return (axis * angle[:, None]).astype(np.float32)


def config0_params(seed: int = 0, count: int = 10_000, sh_degree: int = 0, higher_band_sigma: float = 0.05) -> np.ndarray:
    """10k Gaussians: means U[-1,1]^3, log-scales N(-4, 0.5), axis-angle with random axis and
    angle U(0, pi), opacity logit N(0, 1), SH DC from RGB U[0, 1] (higher bands N(0, sigma)
    when sh_degree > 0)."""
    rng = np.random.default_rng(seed)
    p = empty_params(count)
    p[0:3] = rng.uniform(-1.0, 1.0, size=(3, count))
    p[3:6] = rng.normal(-4.0, 0.5, size=(3, count))
    p[6:9] = random_rotations(rng, count).T
    p[9] = rng.normal(0.0, 1.0, size=count)
    rgb = rng.uniform(0.0, 1.0, size=(3, count))
    for c in range(3):
        p[10 + 16 * c] = (rgb[c] - 0.5) / SH_C0
        if sh_degree > 0:
            p[11 + 16 * c: 26 + 16 * c] = rng.normal(0.0, higher_band_sigma, size=(15, count))
    return p


def perturb(params: np.ndarray, seed: int, sigma: float = 0.01) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (params + rng.normal(0.0, sigma, size=params.shape)).astype(np.float32)


def look_at(eye: np.ndarray, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> tuple[np.ndarray, np.ndarray]:
    """World-to-camera (R, t) with +z forward, +x right, +y down (image convention)."""
    eye = np.asarray(eye, np.float64)
    forward = np.asarray(target, np.float64) - eye
    forward /= np.linalg.norm(forward)
    up = np.asarray(up, np.float64)
    if abs(np.dot(up, forward)) > 0.99:
        up = np.array([0.0, 0.0, 1.0])
    right = np.cross(forward, up)
    right /= np.linalg.norm(right)
    down = np.cross(forward, right)
    R = np.stack([right, down, forward])
    t = -R @ eye
    return R.astype(np.float32), t.astype(np.float32)


def fov_focal(size: int, fov_degrees: float) -> float:
    return 0.5 * size / math.tan(math.radians(fov_degrees) / 2.0)


def sphere_cameras(count: int, radius: float, seed: int, hemisphere: bool = False):
    """Poses on a sphere (or upper hemisphere) looking at the origin."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(count):
        v = rng.normal(size=3)
        if hemisphere:
            v[1] = -abs(v[1])  # y is down in camera space conventions; keep cameras above the plane
        v /= np.linalg.norm(v)
        poses.append(look_at(radius * v))
    return poses


def blob_scene(count: int, seed: int, blobs: int = 64, sh_degree: int = 3, colour_noise: float = 0.05,
               texture: float = 0.0, scale: float = 1.0) -> np.ndarray:
    """Gaussians sampled from a seeded mixture of anisotropic blobs inside the unit ball (the
    large-scale workload): positions from the mixture, log-scales set later from 3-NN distances,
    opacity logit N(0, 1), SH DC from per-blob colours plus N(0, colour_noise) per Gaussian (the
    scene's fine texture), higher bands N(0, 0.05). texture > 0 adds a seeded colour field of that
    amplitude, a sum of plane waves with wavelengths 0.03-0.2 (structure between the spacing of a
    3M-Gaussian scene and of a 200K-point one); it depends on the seed and the position only, so
    sparse points and a dense scene of the same seed share it. scale multiplies the blob centres
    and axes (the same draws)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.5, 1.5, size=(blobs, 3)) * scale
    axes = rng.uniform(0.05, 0.4, size=(blobs, 3)) * scale
    rotations = [axis_angle_matrix(rng.normal(size=3), rng.uniform(0, math.pi)) for _ in range(blobs)]
    colours = rng.uniform(0.1, 0.9, size=(blobs, 3))
    labels = rng.integers(0, blobs, size=count)
    local = rng.normal(size=(count, 3)) * axes[labels]
    R = np.stack(rotations)[labels]
    positions = np.einsum("nij,nj->ni", R, local) + centres[labels]
    p = empty_params(count)
    p[0:3] = positions.T
    p[3:6] = -4.0
    p[6:9] = random_rotations(rng, count).T
    p[9] = rng.normal(0.0, 1.0, size=count)
    rgb = np.clip(colours[labels] + rng.normal(0.0, colour_noise, size=(count, 3)), 0.0, 1.0)
    for c in range(3):
        p[10 + 16 * c] = (rgb[:, c] - 0.5) / SH_C0
        if sh_degree > 0:
            p[11 + 16 * c: 26 + 16 * c] = rng.normal(0.0, 0.05, size=(15, count))
    if texture > 0.0:
        rgb = np.clip(colours[labels] + texture * texture_field(positions / scale, seed) +
                      rng.normal(0.0, colour_noise, size=(count, 3)), 0.0, 1.0)
        for c in range(3):
            p[10 + 16 * c] = (rgb[:, c] - 0.5) / SH_C0
    return p


def texture_field(positions: np.ndarray, seed: int, waves: int = 12) -> np.ndarray:
    """[count, 3] colour offsets in about [-1, 1]: `waves` plane waves with random directions,
    log-uniform wavelengths in [0.03, 0.2], phases and per-channel amplitudes, from the seed."""
    rng = np.random.default_rng([seed, 7])
    directions = rng.normal(size=(waves, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    wavelengths = np.exp(rng.uniform(np.log(0.03), np.log(0.2), size=waves))
    phases = rng.uniform(0.0, 2.0 * math.pi, size=waves)
    amplitudes = rng.normal(size=(waves, 3)) / math.sqrt(waves)
    field = np.zeros((positions.shape[0], 3))
    for j in range(waves):
        wave = np.sin(2.0 * math.pi * (positions @ directions[j]) / wavelengths[j] + phases[j])
        field += wave[:, None] * amplitudes[j][None, :]
    return field


def shell_plane_scene(count: int, seed: int, sh_degree: int = 3, plane_fraction: float = 0.3,
                      texture: float = 0.0, colour_noise: float = 0.05) -> np.ndarray:
    """Gaussians on a radius-2 spherical shell (0.05 thick) plus a ground plane at y = 2 (y points
    down in the camera convention of sphere_cameras(hemisphere=True), so the cameras look at the
    shell from above the plane), x and z in [-4, 4]: the 500,000-Gaussian GPU configuration.
    Log-scales are set later from 3-NN distances; random rotations, opacity logit N(0, 1), SH DC
    from uniform colours, higher bands N(0, 0.05). texture > 0 replaces the uniform colours by
    0.5 + texture * texture_field(position) + N(0, colour_noise): a surface pattern shared by every
    scene of the same seed whatever its count."""
    rng = np.random.default_rng(seed)
    planar = int(round(count * plane_fraction))
    shell = count - planar
    directions = rng.normal(size=(shell, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    radii = 2.0 + rng.uniform(-0.025, 0.025, size=(shell, 1))
    ground = np.stack([rng.uniform(-4.0, 4.0, planar), 2.0 + rng.normal(0.0, 0.01, planar),
                       rng.uniform(-4.0, 4.0, planar)], axis=1)
    positions = np.concatenate([directions * radii, ground])
    p = empty_params(count)
    p[0:3] = positions.T
    p[3:6] = -4.0
    p[6:9] = random_rotations(rng, count).T
    p[9] = rng.normal(0.0, 1.0, size=count)
    rgb = rng.uniform(0.0, 1.0, size=(count, 3))
    for c in range(3):
        p[10 + 16 * c] = (rgb[:, c] - 0.5) / SH_C0
        if sh_degree > 0:
            p[11 + 16 * c: 26 + 16 * c] = rng.normal(0.0, 0.05, size=(15, count))
    if texture > 0.0:
        rgb = np.clip(0.5 + texture * texture_field(positions, seed) + rng.normal(0.0, colour_noise, size=(count, 3)),
                      0.0, 1.0)
        for c in range(3):
            p[10 + 16 * c] = (rgb[:, c] - 0.5) / SH_C0
    return p


def scales_from_mean_dist2(mean_dist2: np.ndarray) -> np.ndarray:
    """gaussian_splatting.cpp:103-113: log-scale = 0.5 * log(max(d2, 1e-7))."""
    return (0.5 * np.log(np.maximum(mean_dist2.astype(np.float32), np.float32(1e-7)))).astype(np.float32)
