"""The fused gradient pass can also project the updated parameters for the next view
(gsn_gaussian_gradients_adam_project). These tests require it to write what the separate path
(gsn_gaussian_gradients_adam and gsn_density_accumulate, then gsn_project_gaussians of the updated
parameters) writes, bit for bit: parameters, moments, density accumulators, every projection
output of the visible Gaussians, and the tile
lists and image rendered from them; for SH degrees 0-3, with parameter rows on and off the
128-byte line grid. A training sequence across iteration 1000 (SH degree change, density step,
opacity cap, each of which drops the projected view) must match the trainer without fusion.
"""
import numpy as np
import pytest

from helpers import config0_camera

pytestmark = pytest.mark.gpu


def _visible_bits(projection, name, counts):
    values = getattr(projection, name).numpy().reshape(projection.count, -1)
    return values[counts > 0].view(np.uint32)


@pytest.mark.parametrize("count", [3000, 3072])
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_fused_projection_matches_the_separate_path(gsn, scenes, degree, count):
    context = gsn.Context()
    host = scenes.config0_params(seed=40 + degree, count=count, sh_degree=3)
    host[3:6] += np.float32(1.0)
    cam = config0_camera(gsn, 192, 160, distance=3.0)
    next_cam = config0_camera(gsn, 192, 160, fov=50.0, distance=2.6)
    image = gsn.DeviceBuffer((160, 192, 3), np.float32)
    rng = np.random.default_rng(degree)
    image_grad = gsn.DeviceBuffer.from_numpy(rng.normal(0, 1e-3, (160, 192, 3)).astype(np.float32))
    c1, c2 = gsn.adam_corrections(3)
    # Density accumulators part-way through a run: the fused pass adds this view's statistics as
    # gsn_density_accumulate does after the separate pass.
    start_sums = rng.uniform(0.0, 1e-2, (2, count)).astype(np.float32)
    start_counts = rng.integers(0, 50, count).astype(np.uint32)
    results = []
    for fused in (False, True):
        params = gsn.DeviceBuffer.from_numpy(host)
        moments = gsn.Moments(count)
        statistics = gsn.DeviceBuffer.zeros((2, count), np.float32)
        absolute = gsn.DeviceBuffer.zeros((2, count), np.float32)
        sums = gsn.DeviceBuffer.from_numpy(start_sums[0])
        absolute_sums = gsn.DeviceBuffer.from_numpy(start_sums[1])
        observations = gsn.DeviceBuffer.from_numpy(start_counts)
        projection = gsn.render(context, params, count, cam, (0, 0, 0), image, degree)
        out_image = gsn.DeviceBuffer((160, 192, 3), np.float32)
        if fused:
            following = gsn.gaussian_gradients_adam_project(context, projection, params, degree, image, image_grad,
                                                            moments, 1.6e-4, c1, c2, statistics, absolute, next_cam,
                                                            accumulate=(sums, absolute_sums, observations))
            gsn.render_projected(context, following, (0, 0, 0), out_image)
        else:
            gsn.gaussian_gradients_adam(context, projection, params, degree, image, image_grad, moments, 1.6e-4, c1,
                                        c2, statistics, absolute)
            gsn.call("gsn_density_accumulate", context.handle, statistics.ptr, absolute.ptr, count, sums.ptr,
                     absolute_sums.ptr, observations.ptr)
            following = gsn.render(context, params, count, next_cam, (0, 0, 0), out_image, degree)
        context.synchronize()
        counts = np.diff(following.record_offsets.numpy())
        results.append({
            "params": params.numpy().view(np.uint32), "moments": [m.view(np.uint32) for m in moments.numpy()],
            "statistics": statistics.numpy().view(np.uint32), "offsets": following.record_offsets.numpy(),
            "sums": np.stack([sums.numpy(), absolute_sums.numpy()]).view(np.uint32),
            "observations": observations.numpy(),
            "counts": counts, "records": following.record_count,
            "ids": following.tile_gaussian_ids.numpy()[:following.record_count],
            "tile_offsets": following.tile_offsets.numpy(), "image": out_image.numpy().view(np.uint32),
            **{name: _visible_bits(following, name, counts)
               for name in ("means", "conic_opacities", "colors", "depths", "tile_rects")}})
    separate, fused = results
    assert separate["records"] > 0 and (separate["counts"] > 0).mean() > 0.5
    assert (separate["observations"] > start_counts).any()
    for key in ("params", "statistics", "sums", "observations", "offsets", "ids", "tile_offsets", "image", "means",
                "conic_opacities", "colors", "depths", "tile_rects"):
        np.testing.assert_array_equal(fused[key], separate[key], err_msg=key)
    for a, b in zip(fused["moments"], separate["moments"]):
        np.testing.assert_array_equal(a, b)


def test_gradient_stage_timed_by_its_dispatch(gsn, scenes):
    """The gradient pass is timed by events that its own dispatch records (hipExtLaunchKernel):
    one launch per call, a positive duration no longer than the host's bracket of the call and
    its synchronisation, and nothing recorded once timing is off."""
    import time
    context = gsn.Context()
    count = 20_000
    host = scenes.config0_params(seed=7, count=count, sh_degree=3)
    cam = config0_camera(gsn, 192, 160, distance=3.0)
    image = gsn.DeviceBuffer((160, 192, 3), np.float32)
    image_grad = gsn.DeviceBuffer.from_numpy(np.full((160, 192, 3), 1e-3, np.float32))
    params = gsn.DeviceBuffer.from_numpy(host)
    moments = gsn.Moments(count)
    statistics = gsn.DeviceBuffer.zeros((2, count), np.float32)
    absolute = gsn.DeviceBuffer.zeros((2, count), np.float32)
    c1, c2 = gsn.adam_corrections(1)
    try:
        for timed in (True, False):
            projection = gsn.render(context, params, count, cam, (0, 0, 0), image, 3)
            context.synchronize()
            context.set_timing(timed, stages=["gaussian_gradients"])
            begin = time.perf_counter()
            following = gsn.gaussian_gradients_adam_project(context, projection, params, 3, image, image_grad,
                                                            moments, 1.6e-4, c1, c2, statistics, absolute, cam)
            context.synchronize()
            bracket_ms = (time.perf_counter() - begin) * 1e3
            ms, launches = context.stage_times()["gaussian_gradients"]
            if timed:
                assert launches == 1 and 0.0 < ms <= bracket_ms, (ms, launches, bracket_ms)
            else:
                assert launches == 0
            gsn.render_projected(context, following, (0, 0, 0), image)  # the record count comes from the scan
            assert following.record_count > 0
    finally:
        context.set_timing(False)
        context.close()


def test_training_with_and_without_fused_projection_is_identical(gsn):
    import gs_native_amd.trainer as trainer_module
    import gs_native_amd.workloads as workloads
    context = gsn.Context()
    params, cams, targets, extent, _ = workloads.build_workload(gsn, context, 20_000, 320, 240, 8)
    results = []
    for overlap, fuse in ((False, False), (True, True), (False, True), (True, False)):
        trainer = trainer_module.Trainer(context, params, cams, targets, extent, completed=985)
        trainer.overlap_backward = overlap
        trainer.fuse_projection = fuse
        for _ in range(30):  # iteration 1000: SH degree 0 -> 1, density step, opacity cap
            trainer.iteration()
        context.synchronize()
        results.append((trainer.count, trainer.params.numpy().copy(), trainer.moments.numpy()))
    print(f"count after iteration 1015: {results[0][0]}", flush=True)
    for count, values, moments in results[1:]:
        assert count == results[0][0]
        np.testing.assert_array_equal(values.view(np.uint32), results[0][1].view(np.uint32))
        for a, b in zip(moments, results[0][2]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
