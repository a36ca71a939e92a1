"""The allocation-free metrics entry point (gsn_image_metrics_static) against gsn_image_metrics,
through the C ABI: two calls in a row, the second on a smaller image, must return the same bits,
without growing the context's scratch or taking device memory."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _image_pair(rng, width, height):
    image = rng.uniform(-0.1, 1.1, size=(height, width, 3)).astype(np.float32)
    target = np.clip(np.rint(np.clip(image, 0, 1) * 255.0 + rng.normal(0, 8, size=image.shape)), 0, 255)
    return image, target.astype(np.uint8)


def _static_metrics(gsn, context, width, height, image, target, out):
    gsn.call("gsn_image_metrics_static", context.handle, width, height, image.ptr, target.ptr, out.ptr)
    context.synchronize()
    return out.numpy()


def test_static_metrics_match_allocating_entry_point(gsn):
    context = gsn.Context()  # fresh: its scratch slots are still empty
    rng = np.random.default_rng(7)
    out = gsn.DeviceBuffer((4,), np.float64)
    # One call on a tiny image first: the runtime's own set-up of a process's first kernel
    # launches (kernel-argument pools and the like) is not the entry point's allocation.
    tiny_image, tiny_target = _image_pair(rng, 16, 16)
    _static_metrics(gsn, context, 16, 16, gsn.DeviceBuffer.from_numpy(tiny_image),
                    gsn.DeviceBuffer.from_numpy(tiny_target), out)
    results = []
    for width, height in ((1920, 1080), (640, 480)):
        image_host, target_host = _image_pair(rng, width, height)
        image = gsn.DeviceBuffer.from_numpy(image_host)
        target = gsn.DeviceBuffer.from_numpy(target_host)
        scratch_before = context.scratch_bytes()
        free_before, _ = gsn.device_memory_info()
        static = _static_metrics(gsn, context, width, height, image, target, out)
        free_after, _ = gsn.device_memory_info()
        assert context.scratch_bytes() == scratch_before, "the static entry point grew the context scratch"
        assert free_after == free_before, "device memory was taken during the static call"
        reference = (ctypes.c_double * 4)()
        gsn.call("gsn_image_metrics", context.handle, width, height, image.ptr, target.ptr, reference)
        np.testing.assert_array_equal(static, np.array(reference[:]))
        results.append(static)
    # The two images differ, so the second call cannot have returned the first call's values.
    assert not np.array_equal(results[0], results[1])
    # Sanity of the values themselves: L1 of the clamped noise is a few percent, MSE positive.
    for loss, l1, ssim, mse in results:
        assert 0.0 < l1 < 0.5 and -1.0 < ssim < 1.0 and mse > 0.0
        assert abs(loss - (0.8 * l1 + 0.2 * (1.0 - ssim))) < 1e-12
    context.close()


def test_static_metrics_reject_sizes_beyond_the_maximum(gsn, context):
    out = gsn.DeviceBuffer((4,), np.float64)
    dummy = gsn.DeviceBuffer((16,), np.float32)
    status = gsn.lib.gsn_image_metrics_static(context.handle, 8192, 8192, dummy.ptr, dummy.ptr, out.ptr)
    assert status != 0
    assert b"static maximum" in gsn.lib.gsn_last_error()
