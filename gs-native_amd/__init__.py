"""Python binding of libgsn_hip.so, the gfx950 Gaussian-splatting training path.

The library is loaded with ctypes; device memory is handled through the library's own
plumbing (gsn_device_malloc / gsn_memcpy_*), so no other GPU runtime is involved. Import fails
loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

PACKAGE = pathlib.Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE / "libgsn_hip.so"

PARAM_COUNT = 58
POSITION, LOG_SCALE, ROTATION, OPACITY, SH = 0, 3, 6, 9, 10
FLT_MAX = float(np.finfo(np.float32).max)


class GsnError(RuntimeError):
    """A C ABI entry point returned a non-zero status."""

    def __init__(self, function: str, status: int, message: str):
        super().__init__(f"{function} failed with status {status}: {message}")
        self.status = status


class Camera(ctypes.Structure):
    """Byte-identical to gsn_camera / gs_native::CameraParams."""

    _fields_ = [
        ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
        ("R_c_w", ctypes.c_float * 9), ("t_c_w", ctypes.c_float * 3),
        ("near_depth", ctypes.c_float), ("far_depth", ctypes.c_float),
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("pixel_offset", ctypes.c_int64),
        ("tile_count_x", ctypes.c_int), ("tile_count_y", ctypes.c_int), ("tile_offset", ctypes.c_int),
    ]

    @classmethod
    def make(cls, width: int, height: int, fx: float, fy: float, cx: float, cy: float,
             rotation=None, translation=None, near: float = 0.2, far: float = FLT_MAX) -> "Camera":
        camera = cls()
        camera.fx, camera.fy, camera.cx, camera.cy = fx, fy, cx, cy
        rotation = np.eye(3, dtype=np.float32) if rotation is None else np.asarray(rotation, np.float32)
        translation = np.zeros(3, np.float32) if translation is None else np.asarray(translation, np.float32)
        for k, value in enumerate(rotation.reshape(9)):
            camera.R_c_w[k] = float(value)
        for k, value in enumerate(translation.reshape(3)):
            camera.t_c_w[k] = float(value)
        camera.near_depth, camera.far_depth = near, far
        camera.width, camera.height = width, height
        camera.tile_count_x = (width + 15) // 16
        camera.tile_count_y = (height + 15) // 16
        return camera

    @property
    def tile_count(self) -> int:
        return self.tile_count_x * self.tile_count_y

    @property
    def pixels(self) -> int:
        return self.width * self.height

    def rotation(self) -> np.ndarray:
        return np.array(self.R_c_w[:], np.float32).reshape(3, 3)

    def translation(self) -> np.ndarray:
        return np.array(self.t_c_w[:], np.float32)


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_U64 = ctypes.c_uint64
_I64 = ctypes.c_int64
_CAM = ctypes.POINTER(Camera)

_PROTOTYPES = {
    "gsn_context_create": [ctypes.POINTER(_P)],
    "gsn_context_destroy": [_P],
    "gsn_context_synchronize": [_P],
    "gsn_device_synchronize": [],
    "gsn_context_stream": [_P],
    "gsn_context_scratch_bytes": [_P],
    "gsn_status_string": [_I],
    "gsn_last_error": [],
    "gsn_device_count": [ctypes.POINTER(_I)],
    "gsn_device_name": [ctypes.c_char_p, ctypes.c_size_t],
    "gsn_device_malloc": [ctypes.POINTER(_P), ctypes.c_size_t],
    "gsn_device_free": [_P],
    "gsn_memcpy_to_device": [_P, _P, ctypes.c_size_t],
    "gsn_memcpy_to_host": [_P, _P, ctypes.c_size_t],
    "gsn_memcpy_device": [_P, _P, ctypes.c_size_t],
    "gsn_memset_device": [_P, _I, ctypes.c_size_t],
    "gsn_device_memory_info": [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    "gsn_event_create": [ctypes.POINTER(_P)],
    "gsn_event_destroy": [_P],
    "gsn_event_record": [_P, _P],
    "gsn_event_elapsed_ms": [_P, _P, ctypes.POINTER(_F)],
    "gsn_project_gaussians": [_P, _P, _I, _CAM, _I, _P, _P, _P, _P, _P, _P],
    "gsn_scan_record_offsets": [_P, _P, _I, ctypes.POINTER(_I)],
    "gsn_build_tile_lists": [_P, _CAM, _I, _P, _P, _P, _I, _P, _P, _P],
    "gsn_render_tiles": [_P, _CAM, _P, _P, _P, _P, _P, ctypes.POINTER(_F), _P],
    "gsn_image_gradients": [_P, _I, _I, _P, _P, _P],
    "gsn_image_metrics": [_P, _I, _I, _P, _P, ctypes.POINTER(ctypes.c_double)],
    "gsn_image_metrics_static": [_P, _I, _I, _P, _P, _P],
    "gsn_image_quantized_error": [_P, _I, _I, _P, _P, _P, ctypes.POINTER(ctypes.c_double)],
    "gsn_gaussian_gradients": [_P, _CAM, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "gsn_gaussian_gradients_adam": [_P, _CAM, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _F, _F, _F, _P, _P],
    "gsn_gaussian_gradients_adam_project": [_P, _CAM, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                                            _F, _F, _F, _P, _P, _CAM, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "gsn_adam_update": [_P, _I, _P, _P, _P, _F, _F, _F],
    "gsn_remap_history": [_P, _P, _I, _I, _P, _P],
    "gsn_zero_rows": [_P, _P, _I, _I, _I],
    "gsn_moment_floats": [_I],
    "gsn_moments_unpack": [_P, _I, _P, _P, _P],
    "gsn_moments_pack": [_P, _I, _P, _P, _P],
    "gsn_density_count": [_P, _P, _I, _P, _P, _P, _F, _F, _F, _F, _I, _F, _P, _P, ctypes.POINTER(ctypes.c_int32)],
    "gsn_density_write": [_P, _P, _I, _P, _P, ctypes.POINTER(ctypes.c_int32), _U64, _U64, _P, _I, _P, _P, _P],
    "gsn_density_normals": [_P, _U64, _U64, _I64, _P],
    "gsn_density_accumulate": [_P, _P, _P, _I, _P, _P, _P],
    "gsn_density_accumulate_observations": [_P, _P, _P, _P, _I, _I, _P, _P],
    "gsn_density_reset_opacity": [_P, _P, _I, _I, _F],
    "gsn_density_clear": [_P, _P, _P, _P, _I],
    "gsn_knn3": [_P, _P, _I, _P, _P],
    "gsn_density_count_proposed": [_P, _P, _I, _P, _P, _P, _P, _F, _F, _F, _I, _P, _P,
                                   ctypes.POINTER(ctypes.c_int32)],
    "gsn_density_features_clear": [_P, _P, _I],
    "gsn_density_features_accumulate": [_P, _P, _P, _P, _P, _I, _P, _I],
    "gsn_density_features": [_P, _P, _I, _P, _P, _P, _P, _I, _F, _F, _F, _P],
    "gsn_density_outcomes": [_P, _P, _P, ctypes.POINTER(ctypes.c_int32), _I, _P, _P, _I, _F, _P],
    "gsn_render_tiles_stored": [_P, _CAM, _P, _P, _P, _P, _P, ctypes.POINTER(_F), _P, _P, _P],
    "gsn_gaussian_gradients_stored": [_P, _CAM, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                                      ctypes.POINTER(_F), _P, _P, _P, _P, _P],
    "gsn_gaussian_gradients_adam_stored": [_P, _CAM, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P,
                                           ctypes.POINTER(_F), _P, _P, _F, _F, _F, _P, _P],
    "gsn_mlp_sizes": [ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "gsn_mlp_sync_weights": [_P, _P, _P],
    "gsn_mlp_forward": [_P, _P, _I, _P, _P, _P],
    "gsn_mlp_train_step": [_P, _P, _P, _P, _I, _P, _P, _P, _P, _F, _I, _P, ctypes.POINTER(ctypes.c_double)],
    "gsn_selftest_wave_sum12": [_P, _P, _P],
    "gsn_selftest_band_mask": [_P, _P, _P, _I, _P],
    "gsn_selftest_sort_pairs": [_P, _P, _P, _I, _I, _P, _P],
    "gsn_context_set_timing": [_P, _I],
    "gsn_context_set_timing_mask": [_P, ctypes.c_uint],
    "gsn_context_stage_times": [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I), _I],
    "gsn_context_count_raster_pairs": [_P, _I],
    "gsn_context_raster_pairs": [_P, ctypes.POINTER(ctypes.c_ulonglong)],
    "gsn_context_prefetch_backward": [_P],
    "gsn_context_prefetch_reused": [_P, ctypes.POINTER(ctypes.c_ulonglong)],
}

STAGES = ("project", "scan", "depth_sort", "emit", "tile_sort", "ranges", "raster_forward", "image_gradients",
          "raster_backward", "gaussian_gradients", "adam", "density", "knn", "accumulate")

_RESTYPES = {
    "gsn_context_destroy": None,
    "gsn_context_stream": _P,
    "gsn_context_scratch_bytes": ctypes.c_size_t,
    "gsn_moment_floats": ctypes.c_size_t,
    "gsn_status_string": ctypes.c_char_p,
    "gsn_last_error": ctypes.c_char_p,
}


def load_library(path: pathlib.Path = LIBRARY_PATH) -> ctypes.CDLL:
    if not path.exists():
        raise ImportError(f"{path} is not built: run __graft_entry__.build() (there is no CPU fallback)")
    library = ctypes.CDLL(str(path))
    for name, argtypes in _PROTOTYPES.items():
        function = getattr(library, name)
        function.argtypes = argtypes
        function.restype = _RESTYPES.get(name, _I)
    return library


lib = load_library()


def exported_symbols() -> list[str]:
    return sorted(_PROTOTYPES)


def check(function: str, status: int) -> None:
    if status != 0:
        raise GsnError(function, status, lib.gsn_last_error().decode(errors="replace"))


def call(function: str, *args) -> None:
    check(function, getattr(lib, function)(*args))


def device_count() -> int:
    count = ctypes.c_int(0)
    call("gsn_device_count", ctypes.byref(count))
    return count.value


def device_name() -> str:
    buffer = ctypes.create_string_buffer(256)
    call("gsn_device_name", buffer, 256)
    return buffer.value.decode()


def device_memory_info() -> tuple[int, int]:
    """(free, total) bytes of the current device."""
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    call("gsn_device_memory_info", ctypes.byref(free), ctypes.byref(total))
    return free.value, total.value


class DeviceBuffer:
    """Owning device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._pointer = ctypes.c_void_p(None)
        if self.nbytes:
            call("gsn_device_malloc", ctypes.byref(self._pointer), self.nbytes)

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._pointer

    @classmethod
    def from_numpy(cls, array: np.ndarray) -> "DeviceBuffer":
        array = np.ascontiguousarray(array)
        buffer = cls(array.shape, array.dtype)
        buffer.upload(array)
        return buffer

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceBuffer":
        buffer = cls(shape, dtype)
        if buffer.nbytes:
            call("gsn_memset_device", buffer.ptr, 0, buffer.nbytes)
        return buffer

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array, dtype=self.dtype)
        if array.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {array.nbytes} != {self.nbytes}")
        if self.nbytes:
            call("gsn_memcpy_to_device", self.ptr, array.ctypes.data_as(ctypes.c_void_p), self.nbytes)

    def numpy(self, count: int | None = None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("gsn_memcpy_to_host", out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes)
        return out

    def view(self, shape, dtype=None) -> "DeviceBuffer":
        """A buffer of `shape` over the start of this allocation (no copy, no ownership): lets a
        capacity-sized allocation serve a smaller logical array, so that counts that change by a
        little (density control) do not reallocate."""
        return _DeviceView(self, shape, self.dtype if dtype is None else dtype)

    def free(self) -> None:
        if self._pointer.value:
            lib.gsn_device_free(self._pointer)
            self._pointer = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DeviceView(DeviceBuffer):
    def __init__(self, owner: DeviceBuffer, shape, dtype):
        self.shape = tuple(int(v) for v in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        if self.nbytes > owner.nbytes:
            raise ValueError(f"a view of {self.nbytes} bytes exceeds its allocation of {owner.nbytes}")
        self._owner = owner
        self._pointer = owner.ptr

    def free(self) -> None:
        self._pointer = ctypes.c_void_p(None)
        self._owner = None


class CapacityBuffer:
    """An allocation that grows by 25% when a larger array is asked for and otherwise hands out
    views of the requested shape."""

    def __init__(self, dtype):
        self.dtype = np.dtype(dtype)
        self.storage = None

    def get(self, shape, zero: bool = False) -> DeviceBuffer:
        shape = tuple(int(v) for v in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        needed = max(int(np.prod(shape, dtype=np.int64)), 1)
        if self.storage is None or self.storage.shape[0] < needed:
            self.storage = None
            self.storage = DeviceBuffer((needed + needed // 4,), self.dtype)
        view = self.storage.view(shape)
        if zero and view.nbytes:
            call("gsn_memset_device", view.ptr, 0, view.nbytes)
        return view


class Context:
    """A HIP stream plus the scratch memory of the kernels."""

    def __init__(self):
        self._handle = ctypes.c_void_p(None)
        call("gsn_context_create", ctypes.byref(self._handle))

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def synchronize(self) -> None:
        call("gsn_context_synchronize", self._handle)

    def scratch_bytes(self) -> int:
        return int(lib.gsn_context_scratch_bytes(self._handle))

    def set_timing(self, enabled: bool, stages=None) -> None:
        """Times every stage (or only the named `stages`) until disabled."""
        if stages is None:
            call("gsn_context_set_timing", self._handle, int(enabled))
            return
        mask = 0
        for name in stages:
            mask |= 1 << STAGES.index(name)
        call("gsn_context_set_timing_mask", self._handle, mask if enabled else 0)

    def stage_times(self) -> dict:
        """{stage: (total milliseconds, timed launches)} since timing was enabled or last read."""
        ms = (ctypes.c_double * len(STAGES))()
        launches = (ctypes.c_int * len(STAGES))()
        call("gsn_context_stage_times", self._handle, ms, launches, len(STAGES))
        return {name: (ms[k], launches[k]) for k, name in enumerate(STAGES)}

    def count_raster_pairs(self, enabled: bool) -> None:
        """While enabled, the raster forward counts the (pixel, splat) pairs it evaluates."""
        call("gsn_context_count_raster_pairs", self._handle, int(enabled))

    def raster_pairs(self) -> int:
        """Evaluated (pixel, splat) pairs since the last read (clears the counter)."""
        value = ctypes.c_ulonglong(0)
        call("gsn_context_raster_pairs", self._handle, ctypes.byref(value))
        return int(value.value)

    def prefetch_backward(self) -> None:
        """Arms the next tile render of the last projection to queue the backward's preparation on
        a second stream (overlapped with the forward raster and the image gradient)."""
        call("gsn_context_prefetch_backward", self._handle)

    def prefetch_reused(self) -> int:
        """Gradient calls on this context that used work queued by an armed render."""
        value = ctypes.c_ulonglong(0)
        call("gsn_context_prefetch_reused", self._handle, ctypes.byref(value))
        return int(value.value)

    def close(self) -> None:
        if self._handle.value:
            lib.gsn_context_destroy(self._handle)
            self._handle = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self._handle = ctypes.c_void_p(None)
        call("gsn_event_create", ctypes.byref(self._handle))

    def record(self, context: Context) -> None:
        call("gsn_event_record", context.handle, self._handle)

    def elapsed_ms(self, end: "Event") -> float:
        value = ctypes.c_float(0.0)
        call("gsn_event_elapsed_ms", self._handle, end._handle, ctypes.byref(value))
        return float(value.value)

    def __del__(self):
        try:
            if self._handle.value:
                lib.gsn_event_destroy(self._handle)
        except Exception:
            pass


def _p(buffer) -> ctypes.c_void_p:
    if buffer is None:
        return ctypes.c_void_p(None)
    if isinstance(buffer, (DeviceBuffer, Moments)):
        return buffer.ptr
    return ctypes.c_void_p(int(buffer))


class Projection:
    """Device outputs of projection + tile lists for one view (the Renderer's getters). The
    per-Gaussian buffers are views of capacity allocations, so resize() to a nearby count (after
    density control) reuses them."""

    def __init__(self, count: int, camera: Camera):
        self.camera = camera
        self._stores = {name: CapacityBuffer(dtype) for name, dtype in (
            ("means", np.float32), ("conic_opacities", np.float32), ("colors", np.float32), ("depths", np.float32),
            ("tile_rects", np.int32), ("record_offsets", np.int32))}
        self.resize(count)
        self.tile_offsets = DeviceBuffer((camera.tile_count + 1,), np.int64)
        self.record_count = 0
        self.tile_gaussian_ids = DeviceBuffer((1,), np.int32)
        self.sorted_keys = None

    def resize(self, count: int) -> None:
        self.count = count
        n = max(count, 1)
        self.means = self._stores["means"].get((n, 2))
        self.conic_opacities = self._stores["conic_opacities"].get((n, 4))
        self.colors = self._stores["colors"].get((n, 3))
        self.depths = self._stores["depths"].get((n,))
        self.tile_rects = self._stores["tile_rects"].get((n, 4))
        self.record_offsets = self._stores["record_offsets"].get((count + 1,))


def render(context: Context, params: DeviceBuffer, count: int, camera: Camera, background, image: DeviceBuffer,
           sh_degree: int, projection: Projection | None = None, keep_keys: bool = False,
           stored_state: tuple | None = None) -> Projection:
    """Projection -> record scan -> tile lists -> raster forward (Renderer::Render).
    stored_state = (final_transmittance, last_contributor) selects the forward that also writes the
    per-pixel state of the stored-state backward."""
    if projection is None:
        projection = Projection(count, camera)
    elif projection.count != count:
        projection.resize(count)
    projection.camera = camera
    if projection.tile_offsets.shape[0] < camera.tile_count + 1:
        projection.tile_offsets = DeviceBuffer((camera.tile_count + 1,), np.int64)
    h = context.handle
    call("gsn_project_gaussians", h, _p(params), count, ctypes.byref(camera), sh_degree, _p(projection.means),
         _p(projection.conic_opacities), _p(projection.colors), _p(projection.depths), _p(projection.tile_rects),
         _p(projection.record_offsets))
    return render_projected(context, projection, background, image, keep_keys, stored_state)


def render_projected(context: Context, projection: Projection, background, image: DeviceBuffer,
                     keep_keys: bool = False, stored_state: tuple | None = None) -> Projection:
    """The rest of render() for a projection whose per-Gaussian outputs are written (by
    gsn_project_gaussians or by the fused gradient pass): record scan -> tile lists -> raster
    forward."""
    h = context.handle
    camera = projection.camera
    count = projection.count
    if projection.tile_offsets.shape[0] < camera.tile_count + 1:
        projection.tile_offsets = DeviceBuffer((camera.tile_count + 1,), np.int64)
    records = ctypes.c_int(0)
    call("gsn_scan_record_offsets", h, _p(projection.record_offsets), count, ctypes.byref(records))
    projection.record_count = records.value
    if projection.tile_gaussian_ids.shape[0] < max(records.value, 1):
        projection.tile_gaussian_ids = DeviceBuffer((int(records.value * 1.25) + 1,), np.int32)
    keys = None
    if keep_keys:
        projection.sorted_keys = DeviceBuffer((max(records.value, 1),), np.uint64)
        keys = projection.sorted_keys
    call("gsn_build_tile_lists", h, ctypes.byref(camera), count, _p(projection.tile_rects), _p(projection.depths),
         _p(projection.record_offsets), records.value, _p(projection.tile_gaussian_ids), _p(projection.tile_offsets),
         _p(keys))
    bg = (ctypes.c_float * 3)(*[float(v) for v in background])
    if stored_state is not None:
        call("gsn_render_tiles_stored", h, ctypes.byref(camera), _p(projection.tile_gaussian_ids),
             _p(projection.tile_offsets), _p(projection.means), _p(projection.conic_opacities), _p(projection.colors),
             bg, _p(image), _p(stored_state[0]), _p(stored_state[1]))
    else:
        call("gsn_render_tiles", h, ctypes.byref(camera), _p(projection.tile_gaussian_ids),
             _p(projection.tile_offsets), _p(projection.means), _p(projection.conic_opacities), _p(projection.colors),
             bg, _p(image))
    return projection


def image_gradients(context: Context, camera: Camera, image: DeviceBuffer, target: DeviceBuffer,
                    out: DeviceBuffer) -> None:
    call("gsn_image_gradients", context.handle, camera.width, camera.height, _p(image), _p(target), _p(out))


def image_metrics(context: Context, width: int, height: int, image: DeviceBuffer, target: DeviceBuffer) -> dict:
    out = (ctypes.c_double * 4)()
    call("gsn_image_metrics", context.handle, width, height, _p(image), _p(target), out)
    mse = out[3]
    psnr = float("inf") if mse <= 0 else 10.0 * np.log10(1.0 / mse)
    return {"loss": out[0], "l1": out[1], "ssim": out[2], "mse": mse, "psnr": psnr}


def quantized_error(context: Context, width: int, height: int, image: DeviceBuffer, target: DeviceBuffer,
                    quantized: DeviceBuffer | None = None) -> tuple[float, float]:
    """(sum of squared 8-bit differences, PSNR dB) of the clamped, half-to-even quantised image
    against the target: the held-out score of the reference's render program."""
    out = (ctypes.c_double * 2)()
    call("gsn_image_quantized_error", context.handle, width, height, _p(image), _p(target), _p(quantized), out)
    return out[0], out[1]


def gaussian_gradients(context: Context, projection: Projection, params: DeviceBuffer, sh_degree: int,
                       image: DeviceBuffer, image_grad: DeviceBuffer, gradients: DeviceBuffer | None,
                       statistics: DeviceBuffer, absolute_statistics: DeviceBuffer,
                       screen: DeviceBuffer | None = None) -> None:
    p = projection
    call("gsn_gaussian_gradients", context.handle, ctypes.byref(p.camera), p.count, sh_degree, _p(params),
         _p(p.tile_gaussian_ids), _p(p.tile_offsets), p.record_count, _p(p.record_offsets), _p(p.tile_rects),
         _p(p.means), _p(p.conic_opacities), _p(p.colors), _p(image), _p(image_grad), _p(gradients), _p(screen),
         _p(statistics), _p(absolute_statistics))


class Moments:
    """Adam moments of `count` Gaussians in the library's blocked layout (gsn_moment_floats
    floats: per 64 Gaussians and row, 64 (first, second) pairs). numpy() / from_numpy() convert
    to and from two [58][count] arrays on the device. The layout of a Gaussian does not depend on
    the count, so a buffer allocated for a larger count serves a smaller one (store)."""

    def __init__(self, count: int, store: "CapacityBuffer | None" = None, zero: bool = True):
        self.count = int(count)
        shape = (max(int(lib.gsn_moment_floats(self.count)), 1),)
        self.buffer = store.get(shape, zero=zero) if store is not None else (
            DeviceBuffer.zeros(shape, np.float32) if zero else DeviceBuffer(shape, np.float32))

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self.buffer.ptr

    @classmethod
    def from_numpy(cls, first: np.ndarray, second: np.ndarray) -> "Moments":
        first = np.ascontiguousarray(first, np.float32)
        moments = cls(first.shape[1])
        if moments.count:
            first_device = DeviceBuffer.from_numpy(first)
            second_device = DeviceBuffer.from_numpy(np.ascontiguousarray(second, np.float32))
            call("gsn_device_synchronize")
            call("gsn_moments_pack", _default_context().handle, moments.count, first_device.ptr, second_device.ptr,
                 moments.ptr)
            _default_context().synchronize()
        return moments

    def numpy(self) -> tuple[np.ndarray, np.ndarray]:
        first = DeviceBuffer((PARAM_COUNT, max(self.count, 1)), np.float32)
        second = DeviceBuffer((PARAM_COUNT, max(self.count, 1)), np.float32)
        if self.count:
            call("gsn_device_synchronize")  # the moments may still be written on another context's stream
            call("gsn_moments_unpack", _default_context().handle, self.count, self.ptr, first.ptr, second.ptr)
            _default_context().synchronize()
        return first.numpy()[:, : self.count], second.numpy()[:, : self.count]


_CONVERSION_CONTEXT = []


def _default_context() -> "Context":
    """A context of its own for moment conversions (synchronised after each use)."""
    if not _CONVERSION_CONTEXT:
        _CONVERSION_CONTEXT.append(Context())
    return _CONVERSION_CONTEXT[0]


def gaussian_gradients_adam(context: Context, projection: Projection, params: DeviceBuffer, sh_degree: int,
                            image: DeviceBuffer, image_grad: DeviceBuffer, moments: Moments,
                            position_lr: float, first_correction: float, second_correction: float,
                            statistics: DeviceBuffer, absolute_statistics: DeviceBuffer) -> None:
    p = projection
    call("gsn_gaussian_gradients_adam", context.handle, ctypes.byref(p.camera), p.count, sh_degree, _p(params),
         _p(p.tile_gaussian_ids), _p(p.tile_offsets), p.record_count, _p(p.record_offsets), _p(p.tile_rects),
         _p(p.means), _p(p.conic_opacities), _p(p.colors), _p(image), _p(image_grad), _p(moments),
         position_lr, first_correction, second_correction, _p(statistics), _p(absolute_statistics))


def gaussian_gradients_adam_project(context: Context, projection: Projection, params: DeviceBuffer, sh_degree: int,
                                    image: DeviceBuffer, image_grad: DeviceBuffer, moments: Moments,
                                    position_lr: float, first_correction: float, second_correction: float,
                                    statistics: DeviceBuffer, absolute_statistics: DeviceBuffer, next_camera: Camera,
                                    next_projection: Projection | None = None,
                                    accumulate: tuple | None = None) -> Projection:
    """gaussian_gradients_adam, then the projection of the updated parameters for next_camera,
    fused (see gsn_gaussian_gradients_adam_project). Returns next_projection (allocated when None)
    with its per-Gaussian outputs written; finish the render with render_projected.
    accumulate: None, or (gradient_sums, absolute_gradient_sums, observation_counts), to which the
    pass adds the statistics as gsn_density_accumulate would (then do not call it for this view)."""
    p = projection
    if next_projection is None or next_projection is projection:
        next_projection = Projection(p.count, next_camera)
    elif next_projection.count != p.count:
        next_projection.resize(p.count)
    next_projection.camera = next_camera
    q = next_projection
    call("gsn_gaussian_gradients_adam_project", context.handle, ctypes.byref(p.camera), p.count, sh_degree,
         _p(params), _p(p.tile_gaussian_ids), _p(p.tile_offsets), p.record_count, _p(p.record_offsets),
         _p(p.tile_rects), _p(p.means), _p(p.conic_opacities), _p(p.colors), _p(image), _p(image_grad), _p(moments),
         position_lr, first_correction, second_correction, _p(statistics), _p(absolute_statistics),
         ctypes.byref(next_camera), _p(q.means), _p(q.conic_opacities), _p(q.colors), _p(q.depths), _p(q.tile_rects),
         _p(q.record_offsets), *(_p(b) for b in (accumulate or (None, None, None))))
    return next_projection


def adam_corrections(step: int) -> tuple[float, float]:
    """Bias corrections in FP64, rounded to FP32 (reference solver.cu:912-914)."""
    first = np.float32(1.0 / (1.0 - 0.9 ** step))
    second = np.float32(1.0 / (1.0 - 0.999 ** step))
    return float(first), float(second)


def adam_update(context: Context, count: int, gradients: DeviceBuffer, params: DeviceBuffer, moments: Moments,
                position_lr: float, step: int) -> None:
    c1, c2 = adam_corrections(step)
    call("gsn_adam_update", context.handle, count, _p(gradients), _p(params), _p(moments), position_lr, c1, c2)


def knn3(context: Context, points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = points.shape[0]
    if n == 0:
        return np.zeros((0, 3), np.int32), np.zeros(0, np.float32)
    device_points = DeviceBuffer.from_numpy(points)
    neighbors = DeviceBuffer((n, 3), np.int32)
    distances = DeviceBuffer((n,), np.float32)
    call("gsn_knn3", context.handle, _p(device_points), n, _p(neighbors), _p(distances))
    context.synchronize()
    return neighbors.numpy(), distances.numpy()


class ProposalNetwork:
    """Device state of the 16 -> 64 -> 64 -> 3 densification proposal network: FP32 master
    parameters, Adam moments and the BF16 copy read by the MFMA kernels."""

    def __init__(self, context: Context, params: np.ndarray):
        sizes = (ctypes.c_int(), ctypes.c_int())
        call("gsn_mlp_sizes", ctypes.byref(sizes[0]), ctypes.byref(sizes[1]))
        self.parameter_count, self.weight_count = sizes[0].value, sizes[1].value
        params = np.ascontiguousarray(params, np.float32)
        if params.shape != (self.parameter_count,):
            raise ValueError(f"expected {self.parameter_count} parameters, got {params.shape}")
        self.context = context
        self.params = DeviceBuffer.from_numpy(params)
        self.first = DeviceBuffer.from_numpy(np.zeros_like(params))
        self.second = DeviceBuffer.from_numpy(np.zeros_like(params))
        self.weights = DeviceBuffer((self.weight_count,), np.uint16)
        self.step = 0
        call("gsn_mlp_sync_weights", context.handle, _p(self.params), _p(self.weights))
        # The network is then ready for use on any context's stream.
        context.synchronize()

    def forward(self, features: DeviceBuffer, count: int, logits: DeviceBuffer, context: "Context | None" = None) -> None:
        """Logits of `count` feature rows, computed on `context`'s stream (the one that produced the
        features and will read the logits; default: the network's own context)."""
        stream = (context or self.context).handle
        call("gsn_mlp_forward", stream, _p(features), count, _p(self.weights), _p(self.params), _p(logits))

    def train_step(self, features: DeviceBuffer, labels: DeviceBuffer, count: int, learning_rate: float,
                   sample_weights: DeviceBuffer | None = None, gradients: DeviceBuffer | None = None,
                   want_loss: bool = False) -> float | None:
        self.step += 1
        loss = ctypes.c_double(0.0)
        call("gsn_mlp_train_step", self.context.handle, _p(features), _p(labels), _p(sample_weights), count,
             _p(self.params), _p(self.first), _p(self.second), _p(self.weights), learning_rate, self.step,
             _p(gradients), ctypes.byref(loss) if want_loss else None)
        return loss.value if want_loss else None


def render_stored(context: Context, projection: Projection, image: DeviceBuffer, background,
                  final_transmittance: DeviceBuffer, last_contributor: DeviceBuffer) -> None:
    """Re-runs the raster forward of an existing projection with the per-pixel state the
    stored-state backward needs (final transmittance, last contributor)."""
    p = projection
    bg = (ctypes.c_float * 3)(*background)
    call("gsn_render_tiles_stored", context.handle, ctypes.byref(p.camera), _p(p.tile_gaussian_ids),
         _p(p.tile_offsets), _p(p.means), _p(p.conic_opacities), _p(p.colors), bg, _p(image),
         _p(final_transmittance), _p(last_contributor))


def gaussian_gradients_stored(context: Context, projection: Projection, params: DeviceBuffer, sh_degree: int,
                              final_transmittance: DeviceBuffer, last_contributor: DeviceBuffer, background,
                              image_grad: DeviceBuffer, gradients: DeviceBuffer | None, statistics: DeviceBuffer,
                              absolute_statistics: DeviceBuffer, screen: DeviceBuffer | None = None) -> None:
    p = projection
    bg = (ctypes.c_float * 3)(*background)
    call("gsn_gaussian_gradients_stored", context.handle, ctypes.byref(p.camera), p.count, sh_degree, _p(params),
         _p(p.tile_gaussian_ids), _p(p.tile_offsets), p.record_count, _p(p.record_offsets), _p(p.tile_rects),
         _p(p.means), _p(p.conic_opacities), _p(p.colors), _p(final_transmittance), _p(last_contributor), bg,
         _p(image_grad), _p(gradients), _p(screen), _p(statistics), _p(absolute_statistics))
