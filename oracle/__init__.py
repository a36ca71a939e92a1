"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker of the HIP path. The product never imports it.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "liboracle.so"


def build() -> pathlib.Path:
    subprocess.run(["sh", str(HERE / "build.sh")], check=True)
    return LIBRARY_PATH


class Camera(ctypes.Structure):
    _fields_ = [
        ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
        ("R_c_w", ctypes.c_float * 9), ("t_c_w", ctypes.c_float * 3),
        ("near_depth", ctypes.c_float), ("far_depth", ctypes.c_float),
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("pixel_offset", ctypes.c_int64),
        ("tile_count_x", ctypes.c_int), ("tile_count_y", ctypes.c_int), ("tile_offset", ctypes.c_int),
    ]


def as_camera(camera) -> Camera:
    """Copies any ctypes camera with the same layout."""
    out = Camera()
    ctypes.memmove(ctypes.byref(out), ctypes.byref(camera), ctypes.sizeof(Camera))
    return out


def _load() -> ctypes.CDLL:
    if not LIBRARY_PATH.exists():
        build()
    library = ctypes.CDLL(str(LIBRARY_PATH))
    P = ctypes.c_void_p
    I = ctypes.c_int
    F = ctypes.c_float
    CAM = ctypes.POINTER(Camera)
    signatures = {
        "oracle_set_threads": ([I], None),
        "oracle_rotation_f32": ([F, F, F, P], None),
        "oracle_rotation_f64": ([ctypes.c_double, ctypes.c_double, ctypes.c_double, P], None),
        "oracle_sh_basis_f64": ([ctypes.c_double, ctypes.c_double, ctypes.c_double, P], None),
        "oracle_expf": ([F], F),
        "oracle_logf": ([F], F),
        "oracle_sincosf": ([F, P, P], None),
        "oracle_project_f32": ([P, I, CAM, I, P, P, P, P, P, P], None),
        "oracle_scan_i32": ([P, I], I),
        "oracle_tile_lists": ([CAM, I, P, P, P, I, P, P, P], None),
        "oracle_render_f32": ([CAM, P, P, P, P, P, P, P, P], I),
        "oracle_project_f64": ([P, I, CAM, I, P, P, P, P, P, P], None),
        "oracle_render_f64": ([CAM, I, P, P, P, P, P, P, P, P, P], I),
        "oracle_loss_f64": ([P, P, I, I], ctypes.c_double),
        "oracle_image_gradients_f64": ([P, P, I, I, P], None),
        "oracle_backward_f64": ([CAM, I, P, P, P, P, P, P, P, P, P, P, P, P], None),
        "oracle_map_f64": ([P, I, CAM, I, P, P, P, P, P], None),
        "oracle_adam_f32": ([I, P, P, P, P, F, F, F], None),
        "oracle_remap_f32": ([P, I, I, P, P, P, P], None),
        "oracle_density_count_f32": ([P, I, P, P, P, F, F, F, F, I, F, P, P], None),
        "oracle_density_count_proposed_f32": ([P, I, P, F, F, F, I, P, P], None),
        "oracle_density_write_f32": ([P, I, P, P, ctypes.c_uint64, ctypes.c_uint64, P, I, P, P, P], None),
        "oracle_jitter_normal": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64], F),
        "oracle_philox4x32_10": ([P, ctypes.c_uint32, ctypes.c_uint32], None),
        "oracle_knn3": ([P, I, P, P], None),
    }
    for name, (argtypes, restype) in signatures.items():
        function = getattr(library, name)
        function.argtypes = argtypes
        function.restype = restype
    return library


lib = _load()


def _ptr(array):
    if array is None:
        return ctypes.c_void_p(None)
    return array.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def set_threads(threads: int) -> None:
    """Threads of the oracle's parallel loops (1 = sequential, the default)."""
    lib.oracle_set_threads(int(threads))


def rotation_f32(rotation) -> np.ndarray:
    out = np.zeros(9, np.float32)
    lib.oracle_rotation_f32(*(float(v) for v in rotation), _ptr(out))
    return out.reshape(3, 3)


def rotation_f64(rotation) -> np.ndarray:
    out = np.zeros(9)
    lib.oracle_rotation_f64(*(float(v) for v in rotation), _ptr(out))
    return out.reshape(3, 3)


def sh_basis_f64(direction) -> np.ndarray:
    out = np.zeros(16)
    lib.oracle_sh_basis_f64(*(float(v) for v in direction), _ptr(out))
    return out


def expf(x: float) -> float:
    return lib.oracle_expf(x)


def logf(x: float) -> float:
    return lib.oracle_logf(x)


def sincosf(x: float) -> tuple[float, float]:
    s, c = ctypes.c_float(), ctypes.c_float()
    lib.oracle_sincosf(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def project_f32(params: np.ndarray, camera, sh_degree: int) -> dict:
    params = _f32(params)
    n = params.shape[1]
    out = {
        "means": np.zeros((n, 2), np.float32), "conic_opacities": np.zeros((n, 4), np.float32),
        "colors": np.zeros((n, 3), np.float32), "depths": np.zeros(n, np.float32),
        "tile_rects": np.zeros((n, 4), np.int32), "record_counts": np.zeros(n + 1, np.int32),
    }
    cam = as_camera(camera)
    lib.oracle_project_f32(_ptr(params), n, ctypes.byref(cam), sh_degree, _ptr(out["means"]),
                           _ptr(out["conic_opacities"]), _ptr(out["colors"]), _ptr(out["depths"]),
                           _ptr(out["tile_rects"]), _ptr(out["record_counts"]))
    return out


def tile_lists(camera, projected: dict) -> dict:
    counts = projected["record_counts"].copy()
    n = counts.shape[0] - 1
    total = lib.oracle_scan_i32(_ptr(counts), n)
    cam = as_camera(camera)
    ids = np.zeros(max(total, 1), np.int32)
    keys = np.zeros(max(total, 1), np.uint64)
    offsets = np.zeros(cam.tile_count_x * cam.tile_count_y + 1, np.int64)
    lib.oracle_tile_lists(ctypes.byref(cam), n, _ptr(projected["tile_rects"]), _ptr(projected["depths"]),
                          _ptr(counts), total, _ptr(ids), _ptr(offsets), _ptr(keys))
    return {"record_offsets": counts, "record_count": total, "tile_gaussian_ids": ids[:total],
            "sorted_keys": keys[:total], "tile_offsets": offsets}


def render_f32(camera, projected: dict, lists: dict, background) -> tuple[np.ndarray, np.ndarray]:
    cam = as_camera(camera)
    image = np.zeros((cam.height, cam.width, 3), np.float32)
    ambiguous = np.zeros((cam.height, cam.width), np.uint8)
    bg = _f32(background)
    ids = np.ascontiguousarray(lists["tile_gaussian_ids"], np.int32)
    if ids.size == 0:
        ids = np.zeros(1, np.int32)
    lib.oracle_render_f32(ctypes.byref(cam), _ptr(ids), _ptr(lists["tile_offsets"]), _ptr(projected["means"]),
                          _ptr(projected["conic_opacities"]), _ptr(projected["colors"]), _ptr(bg), _ptr(image),
                          _ptr(ambiguous))
    return image, ambiguous


def project_f64(params: np.ndarray, camera, sh_degree: int) -> dict:
    params = _f32(params)
    n = params.shape[1]
    out = {
        "means": np.zeros((n, 2)), "conics": np.zeros((n, 4)), "colors": np.zeros((n, 3)), "depths": np.zeros(n),
        "tile_rects": np.zeros((n, 4), np.int32), "visible": np.zeros(n, np.uint8),
    }
    cam = as_camera(camera)
    lib.oracle_project_f64(_ptr(params), n, ctypes.byref(cam), sh_degree, _ptr(out["means"]), _ptr(out["conics"]),
                           _ptr(out["colors"]), _ptr(out["depths"]), _ptr(out["tile_rects"]), _ptr(out["visible"]))
    return out


def render_f64(camera, projected: dict, background) -> tuple[np.ndarray, np.ndarray]:
    cam = as_camera(camera)
    n = projected["depths"].shape[0]
    image = np.zeros((cam.height, cam.width, 3))
    ambiguous = np.zeros((cam.height, cam.width), np.uint8)
    bg = _f32(background)
    lib.oracle_render_f64(ctypes.byref(cam), n, _ptr(projected["means"]), _ptr(projected["conics"]),
                          _ptr(projected["colors"]), _ptr(projected["depths"]), _ptr(projected["tile_rects"]),
                          _ptr(projected["visible"]), _ptr(bg), _ptr(image), _ptr(ambiguous))
    return image, ambiguous


def render_reference(params: np.ndarray, camera, background, sh_degree: int) -> np.ndarray:
    """test/cpu_reference.h RenderReference in FP64."""
    return render_f64(camera, project_f64(params, camera, sh_degree), background)[0]


def loss_f64(image: np.ndarray, target: np.ndarray) -> float:
    image = np.ascontiguousarray(image, np.float64)
    target = np.ascontiguousarray(target, np.uint8)
    return lib.oracle_loss_f64(_ptr(image), _ptr(target), image.shape[1], image.shape[0])


def image_gradients_f64(image: np.ndarray, target: np.ndarray) -> np.ndarray:
    image = np.ascontiguousarray(image, np.float64)
    target = np.ascontiguousarray(target, np.uint8)
    out = np.zeros_like(image)
    lib.oracle_image_gradients_f64(_ptr(image), _ptr(target), image.shape[1], image.shape[0], _ptr(out))
    return out


def backward_f64(camera, projected: dict, image: np.ndarray, image_grad: np.ndarray,
                 pixel_flags: np.ndarray | None = None) -> dict:
    cam = as_camera(camera)
    n = projected["depths"].shape[0]
    screen = np.zeros((11, n))
    magnitude = np.zeros((11, n))
    tainted = np.zeros(n, np.uint8)
    image = np.ascontiguousarray(image, np.float64)
    image_grad = np.ascontiguousarray(image_grad, np.float64)
    lib.oracle_backward_f64(ctypes.byref(cam), n, _ptr(projected["means"]), _ptr(projected["conics"]),
                            _ptr(projected["colors"]), _ptr(projected["depths"]), _ptr(projected["tile_rects"]),
                            _ptr(projected["visible"]), _ptr(image), _ptr(image_grad),
                            None if pixel_flags is None else _ptr(np.ascontiguousarray(pixel_flags, np.uint8)),
                            _ptr(screen), _ptr(magnitude), _ptr(tainted))
    return {"screen": screen, "magnitude": magnitude, "tainted": tainted.astype(bool)}


def map_f64(params: np.ndarray, camera, sh_degree: int, screen: np.ndarray, observed: np.ndarray) -> dict:
    params = _f32(params)
    n = params.shape[1]
    cam = as_camera(camera)
    gradients = np.zeros((58, n))
    statistics = np.zeros((2, n))
    absolute = np.zeros((2, n))
    screen = np.ascontiguousarray(screen, np.float64)
    observed = np.ascontiguousarray(observed, np.uint8)
    lib.oracle_map_f64(_ptr(params), n, ctypes.byref(cam), sh_degree, _ptr(screen), _ptr(observed),
                       _ptr(gradients), _ptr(statistics), _ptr(absolute))
    return {"gradients": gradients, "statistics": statistics, "absolute_statistics": absolute}


def gradients_f64(params, camera, sh_degree, background, target, rects_from: dict | None = None) -> dict:
    """Full FP64 chain: project -> render -> dL/dimage -> backward -> map. rects_from (an FP32
    projection) pins the tile rectangles and visibility, so that FP64 and FP32 runs blend the
    same (tile, Gaussian) pairs."""
    projected = project_f64(params, camera, sh_degree)
    if rects_from is not None:
        visible = rects_from["record_counts"][:-1] > 0
        projected["tile_rects"] = np.ascontiguousarray(rects_from["tile_rects"], np.int32)
        projected["visible"] = (visible & (projected["visible"] > 0)).astype(np.uint8)
    image, ambiguous = render_f64(camera, projected, background)
    image_grad = image_gradients_f64(image, target)
    # The L1 term's sign(image - target) is decided within FP32 rendering error (relative ~1e-5
    # of the pixel value) at some pixels; an FP32 run may take the other sign there.
    difference = np.abs(image - np.asarray(target, np.float64) / 255.0)
    loss_ambiguous = np.any((difference <= 1e-5 * np.abs(image)) & (image != 0.0), axis=2)
    back = backward_f64(camera, projected, image, image_grad, loss_ambiguous)
    observed = projected["visible"]
    mapped = map_f64(params, camera, sh_degree, back["screen"], observed)
    return {"projected": projected, "image": image, "ambiguous": ambiguous, "image_grad": image_grad, **back,
            **mapped}


def raster_gradients_f64(camera, projected32: dict, background, target, image_grad=None) -> dict:
    """FP64 render, dL/dimage and replay backward on the FP32 projection's outputs (means, conics,
    colours, depths, rectangles and visibility as the device computed them). Isolates the
    rasteriser from the rounding of the projection, whose accuracy is checked on its own.
    image_grad (optional) replaces the FP64 dL/dimage, e.g. by the device's own, so that the
    comparison also excludes the loss gradient (checked on its own as well)."""
    visible = (projected32["record_counts"][:-1] > 0).astype(np.uint8)
    projected = {
        "means": projected32["means"].astype(np.float64), "conics": projected32["conic_opacities"].astype(np.float64),
        "colors": projected32["colors"].astype(np.float64), "depths": projected32["depths"].astype(np.float64),
        "tile_rects": np.ascontiguousarray(projected32["tile_rects"], np.int32), "visible": visible,
    }
    image, ambiguous = render_f64(camera, projected, background)
    if image_grad is None:
        image_grad = image_gradients_f64(image, target)
        difference = np.abs(image - np.asarray(target, np.float64) / 255.0)
        loss_ambiguous = np.any((difference <= 1e-5 * np.abs(image)) & (image != 0.0), axis=2)
    else:
        image_grad = np.ascontiguousarray(image_grad, np.float64)
        loss_ambiguous = None
    back = backward_f64(camera, projected, image, image_grad, loss_ambiguous)
    return {"image": image, "ambiguous": ambiguous, "image_grad": image_grad, **back}


def adam_f32(gradients, params, first, second, position_lr, first_correction, second_correction):
    gradients, params, first, second = (_f32(a).copy() for a in (gradients, params, first, second))
    n = params.shape[1]
    lib.oracle_adam_f32(n, _ptr(gradients), _ptr(params), _ptr(first), _ptr(second), position_lr, first_correction,
                        second_correction)
    return params, first, second


def remap_f32(parent_ids, previous_first, previous_second, count):
    previous_first = _f32(previous_first)
    previous_second = _f32(previous_second)
    first = np.zeros((58, count), np.float32)
    second = np.zeros((58, count), np.float32)
    parents = None if parent_ids is None else np.ascontiguousarray(parent_ids, np.int32)
    lib.oracle_remap_f32(_ptr(parents), previous_first.shape[1], count, _ptr(previous_first), _ptr(previous_second),
                         _ptr(first), _ptr(second))
    return first, second


def density_count_f32(params, sums, abs_sums, counts, gradient_threshold, split_threshold, min_opacity,
                      max_world_size, prune_large, absolute_gradient_threshold):
    params = _f32(params)
    n = params.shape[1]
    decisions = np.zeros((n + 1, 4), np.int32)
    offsets = np.zeros((n + 1, 4), np.int32)
    lib.oracle_density_count_f32(_ptr(params), n, _ptr(_f32(sums)), _ptr(_f32(abs_sums)),
                                 _ptr(np.ascontiguousarray(counts, np.uint32)), gradient_threshold, split_threshold,
                                 min_opacity, max_world_size, int(prune_large), absolute_gradient_threshold,
                                 _ptr(decisions), _ptr(offsets))
    return decisions, offsets


def density_count_proposed_f32(params, proposal, split_threshold, min_opacity, max_world_size, prune_large):
    params = _f32(params)
    n = params.shape[1]
    decisions = np.zeros((n + 1, 4), np.int32)
    offsets = np.zeros((n + 1, 4), np.int32)
    lib.oracle_density_count_proposed_f32(_ptr(params), n, _ptr(_f32(proposal)), split_threshold, min_opacity,
                                          max_world_size, int(prune_large), _ptr(decisions), _ptr(offsets))
    return decisions, offsets


def density_write_f32(params, decisions, offsets, seed, counter):
    params = _f32(params)
    n = params.shape[1]
    totals = offsets[n]
    out_count = int(totals[0] + totals[1] + 2 * totals[3])
    out = np.zeros((58, max(out_count, 1)), np.float32)
    history = np.zeros(max(out_count, 1), np.int32)
    parents = np.zeros(max(out_count, 1), np.int32)
    flags = np.zeros(max(out_count, 1), np.float32)
    lib.oracle_density_write_f32(_ptr(params), n, _ptr(np.ascontiguousarray(decisions, np.int32)),
                                 _ptr(np.ascontiguousarray(offsets, np.int32)), seed, counter, _ptr(out), out_count,
                                 _ptr(history), _ptr(parents), _ptr(flags))
    return out[:, :out_count], history[:out_count], parents[:out_count], flags[:out_count]


def jitter_normals(seed: int, counter: int, n: int) -> np.ndarray:
    return np.array([lib.oracle_jitter_normal(seed, counter, i) for i in range(n)], np.float32)


def knn3(points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    points = _f32(points).reshape(-1, 3)
    n = points.shape[0]
    neighbors = np.zeros((max(n, 1), 3), np.int32)
    distances = np.zeros(max(n, 1), np.float32)
    lib.oracle_knn3(_ptr(points), n, _ptr(neighbors), _ptr(distances))
    return neighbors[:n], distances[:n]
