"""Renderer parity on the GPU: projection, tile lists and the forward rasteriser against the
oracle (bit-exact integer outputs, FP32 mirror and FP64 reference images), plus the
reference's own renderer cases (test/renderer_test.cu)."""
import ctypes
import json
import pathlib

import numpy as np
import pytest

from helpers import GpuRun, config0_camera, reference_camera

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def test_wave_sum12_lane_layout(gsn, context):
    rng = np.random.default_rng(0)
    values = rng.normal(size=(64, 12)).astype(np.float32)
    source = gsn.DeviceBuffer.from_numpy(values)
    out = gsn.DeviceBuffer((64, 3), np.float32)
    gsn.call("gsn_selftest_wave_sum12", context.handle, source.ptr, out.ptr)
    context.synchronize()
    result = out.numpy()
    totals = values.astype(np.float64).sum(axis=0)
    for lane in range(64):
        row = lane // 16
        np.testing.assert_allclose(result[lane], totals[3 * row: 3 * row + 3], rtol=1e-5, atol=1e-5)


def test_band_mask_is_conservative(gsn, context):
    """The device band mask never drops a band that holds a pixel with alpha >= 1/255 (checked in
    FP64 with the per-pixel rule of the rasteriser, at a threshold 0.1% below 1/255), over random
    ellipses from 0.3 to 400 px in any orientation and opacity, placed around random tiles; and it
    does drop bands (it is not the trivial all-ones mask)."""
    rng = np.random.default_rng(5)
    count = 40000
    s1, s2 = np.exp(rng.uniform(np.log(0.3), np.log(400.0), (2, count)))
    theta = rng.uniform(0.0, np.pi, count)
    cos, sin = np.cos(theta), np.sin(theta)
    xx = cos * cos * s1 * s1 + sin * sin * s2 * s2 + 0.3
    yy = sin * sin * s1 * s1 + cos * cos * s2 * s2 + 0.3
    xy = cos * sin * (s1 * s1 - s2 * s2)
    det = xx * yy - xy * xy
    a, b, c = yy / det, -xy / det, xx / det
    opacity = np.where(rng.random(count) < 0.5, rng.uniform(1 / 255, 0.05, count), rng.uniform(0.05, 1.0, count))
    means = rng.uniform(0.0, 64.0, (count, 2))
    origins = 16 * rng.integers(-3, 8, (count, 2)).astype(np.int32)
    splats = np.stack([means[:, 0], means[:, 1], a, b, c, opacity], axis=1).astype(np.float32)
    masks = gsn.DeviceBuffer((count,), np.int32)
    # The device buffers stay referenced until the kernel has finished.
    device_splats = gsn.DeviceBuffer.from_numpy(splats)
    device_origins = gsn.DeviceBuffer.from_numpy(origins)
    gsn.call("gsn_selftest_band_mask", context.handle, device_splats.ptr, device_origins.ptr, count, masks.ptr)
    context.synchronize()
    del device_splats, device_origins
    masks = masks.numpy()
    s = splats.astype(np.float64)
    columns = np.arange(16.0)[None, :]
    rows = np.arange(4.0)[:, None]
    dropped = 0
    for i in range(count):
        for band in range(4):
            if (masks[i] >> band) & 1:
                continue
            dropped += 1
            dx = s[i, 0] - (origins[i, 0] + columns)
            dy = s[i, 1] - (origins[i, 1] + 4 * band + rows)
            power = -0.5 * (s[i, 2] * dx * dx + s[i, 4] * dy * dy) - s[i, 3] * dx * dy
            alpha = np.minimum(0.99, s[i, 5] * np.exp(power))
            assert not np.any((power <= 0.0) & (alpha >= (1.0 / 255.0) * (1.0 - 1e-3))), (i, band)
    assert dropped > count  # most cases lie away from the tile


def _compare_projection(run, expected):
    actual = run.projected()
    counts = expected["record_counts"][:-1]
    np.testing.assert_array_equal(np.diff(run.projection.record_offsets.numpy()), counts)
    visible = counts > 0
    np.testing.assert_array_equal(actual["tile_rects"][visible], expected["tile_rects"][visible])
    for key in ("means", "conic_opacities", "colors", "depths"):
        a = actual[key][visible].view(np.uint32)
        e = expected[key][visible].view(np.uint32)
        mismatches = np.argwhere(a != e)
        assert mismatches.size == 0, f"{key}: {len(mismatches)} bit mismatches, first {mismatches[:3]}"
    return visible


@pytest.mark.parametrize("sh_degree", [0, 1, 2, 3])
def test_projection_and_tile_lists_bit_exact(gsn, context, oracle, scenes, sh_degree):
    params = scenes.config0_params(seed=7 + sh_degree, count=10_000, sh_degree=3)
    cam = config0_camera(gsn)
    run = GpuRun(gsn, context, params, cam, sh_degree=sh_degree, keep_keys=True)
    expected = oracle.project_f32(params, cam, sh_degree)
    visible = _compare_projection(run, expected)
    assert visible.sum() > 1000
    lists = oracle.tile_lists(cam, expected)
    actual = run.lists()
    assert actual["record_count"] == lists["record_count"]
    np.testing.assert_array_equal(actual["record_offsets"], lists["record_offsets"])
    np.testing.assert_array_equal(actual["tile_offsets"], lists["tile_offsets"])
    np.testing.assert_array_equal(actual["tile_gaussian_ids"], lists["tile_gaussian_ids"])
    np.testing.assert_array_equal(actual["sorted_keys"], lists["sorted_keys"])


def test_tile_lists_large_bit_exact(gsn, context, oracle, scenes):
    """1080p, 200k Gaussians: several radix-sort tiles and a 13-bit tile key."""
    params = scenes.blob_scene(200_000, seed=3, sh_degree=3)
    params[3:6] = np.float32(-3.2)
    R, t = scenes.look_at(np.array([0.3, -0.8, 4.5]))
    f = scenes.fov_focal(1920, 60.0)
    cam = gsn.Camera.make(1920, 1080, f, f, 959.5, 539.5, R, t)
    run = GpuRun(gsn, context, params, cam, sh_degree=3, keep_keys=True)
    expected = oracle.project_f32(params, cam, 3)
    _compare_projection(run, expected)
    lists = oracle.tile_lists(cam, expected)
    actual = run.lists()
    assert actual["record_count"] == lists["record_count"] > 100_000
    np.testing.assert_array_equal(actual["tile_offsets"], lists["tile_offsets"])
    np.testing.assert_array_equal(actual["tile_gaussian_ids"], lists["tile_gaussian_ids"])
    np.testing.assert_array_equal(actual["sorted_keys"], lists["sorted_keys"])


def test_tile_lists_extreme_footprints_bit_exact(gsn, context, oracle, scenes):
    """Records are emitted in 4096-record blocks: blocks owned by thousands of one- and two-tile
    Gaussians, and full-screen Gaussians whose records span two or three blocks."""
    tiny = scenes.blob_scene(40_000, seed=5, sh_degree=0)
    tiny[3:6] = np.float32(-7.5)
    huge = scenes.blob_scene(6, seed=6, sh_degree=0)
    huge[0:3] = np.array([[0.0, 0.3, -0.4, 0.5, 0.2, -0.1], [0.0, -0.2, 0.1, 0.3, -0.3, 0.2],
                          [3.0, 3.2, 2.8, 3.5, 2.5, 3.1]], np.float32)
    huge[3:6] = np.float32(-0.3)
    params = np.concatenate([tiny[:, :20_000], huge, tiny[:, 20_000:]], axis=1)
    R, t = scenes.look_at(np.array([0.3, -0.8, 4.5]))
    f = scenes.fov_focal(1920, 60.0)
    cam = gsn.Camera.make(1920, 1080, f, f, 959.5, 539.5, R, t)
    run = GpuRun(gsn, context, params, cam, sh_degree=0, keep_keys=True)
    expected = oracle.project_f32(params, cam, 0)
    _compare_projection(run, expected)
    lists = oracle.tile_lists(cam, expected)
    actual = run.lists()
    counts = np.diff(lists["record_offsets"])
    assert counts.max() > 4096, "some Gaussian spans two or more blocks"
    assert (counts == 1).sum() > 2 * 4096, "blocks owned by one-record Gaussians"
    assert actual["record_count"] == lists["record_count"]
    np.testing.assert_array_equal(actual["tile_offsets"], lists["tile_offsets"])
    np.testing.assert_array_equal(actual["tile_gaussian_ids"], lists["tile_gaussian_ids"])
    np.testing.assert_array_equal(actual["sorted_keys"], lists["sorted_keys"])


def test_tile_lists_rebuilt_without_queued_depth_order(gsn, context, scenes):
    """render() runs projection -> scan -> tile lists, and the scan queues the depth order that the
    tile-list build reuses. A second build with the same buffers finds no queued order (it is
    consumed) and sorts by depth itself; both builds must agree bit for bit."""
    params = scenes.blob_scene(50_000, seed=9, sh_degree=0)
    params[3:6] = np.float32(-3.5)
    R, t = scenes.look_at(np.array([0.2, -0.5, 4.0]))
    f = scenes.fov_focal(640, 60.0)
    cam = gsn.Camera.make(640, 480, f, f, 319.5, 239.5, R, t)
    run = GpuRun(gsn, context, params, cam, sh_degree=0, keep_keys=True)
    first = run.lists()
    p = run.projection
    r = p.record_count
    ids = gsn.DeviceBuffer((max(r, 1),), np.int32)
    offsets = gsn.DeviceBuffer((cam.tile_count + 1,), np.int64)
    keys = gsn.DeviceBuffer((max(r, 1),), np.uint64)
    gsn.call("gsn_build_tile_lists", context.handle, ctypes.byref(cam), p.count, p.tile_rects.ptr, p.depths.ptr,
             p.record_offsets.ptr, r, ids.ptr, offsets.ptr, keys.ptr)
    context.synchronize()
    assert r > 10_000
    np.testing.assert_array_equal(offsets.numpy(), first["tile_offsets"])
    np.testing.assert_array_equal(ids.numpy()[:r], first["tile_gaussian_ids"])
    np.testing.assert_array_equal(keys.numpy()[:r], first["sorted_keys"])


def test_tile_lists_with_empty_rectangles_between_records(gsn, context):
    """Inputs the projection never writes: 5,000 Gaussians marked visible by their record offsets
    but with empty tile rectangles, between two runs of two-record Gaussians. The first
    4096-record block then has more owners than positions, and the emission finds each record's
    owner by binary search. The lists must equal a stable sort by (tile, depth) of the records in
    depth order, which here is id order (depths increase with the id)."""
    normal_before, empty, normal_after = 2000, 5000, 5000
    count = normal_before + empty + normal_after
    f = 600.0
    cam = gsn.Camera.make(640, 480, f, f, 319.5, 239.5, np.eye(3), np.zeros(3))
    tiles_x, tiles_y = cam.tile_count_x, cam.tile_count_y
    index = np.arange(count)
    x0 = (index * 7) % (tiles_x - 1)
    y0 = (index * 13) % tiles_y
    rects = np.stack([x0, y0, x0 + 2, y0 + 1], axis=1).astype(np.int32)
    hollow = (index >= normal_before) & (index < normal_before + empty)
    rects[hollow] = (5, 5, 5, 5)
    depths = (1.0 + 1e-3 * index).astype(np.float32)
    areas = (rects[:, 2] - rects[:, 0]) * (rects[:, 3] - rects[:, 1])
    record_count = int(areas.sum())
    tiles = np.concatenate([(rects[i, 1] * tiles_x + rects[i, 0] + np.arange(areas[i])) for i in range(count)])
    ids = np.repeat(index, areas)
    order = np.argsort(tiles, kind="stable")
    expected_keys = (tiles[order].astype(np.uint64) << np.uint64(32)) | depths[ids[order]].view(np.uint32).astype(
        np.uint64)
    expected_offsets = np.searchsorted(tiles[order], np.arange(cam.tile_count + 1), side="left").astype(np.int64)

    record_offsets = gsn.DeviceBuffer.from_numpy(np.arange(count + 1, dtype=np.int32))  # every Gaussian visible
    out_ids = gsn.DeviceBuffer((record_count,), np.int32)
    out_offsets = gsn.DeviceBuffer((cam.tile_count + 1,), np.int64)
    out_keys = gsn.DeviceBuffer((record_count,), np.uint64)
    device_rects = gsn.DeviceBuffer.from_numpy(rects)
    device_depths = gsn.DeviceBuffer.from_numpy(depths)
    gsn.call("gsn_build_tile_lists", context.handle, ctypes.byref(cam), count, device_rects.ptr, device_depths.ptr,
             record_offsets.ptr, record_count, out_ids.ptr, out_offsets.ptr, out_keys.ptr)
    context.synchronize()
    assert record_count == 2 * (normal_before + normal_after)
    np.testing.assert_array_equal(out_ids.numpy(), ids[order])
    np.testing.assert_array_equal(out_offsets.numpy(), expected_offsets)
    np.testing.assert_array_equal(out_keys.numpy(), expected_keys)


@pytest.mark.parametrize("sh_degree", [0, 3])
def test_forward_matches_fp32_oracle(gsn, context, oracle, scenes, sh_degree):
    params = scenes.config0_params(seed=11, count=10_000, sh_degree=3)
    cam = config0_camera(gsn)
    background = (0.05, 0.1, 0.15)
    run = GpuRun(gsn, context, params, cam, background, sh_degree)
    projected = oracle.project_f32(params, cam, sh_degree)
    lists = oracle.tile_lists(cam, projected)
    expected, ambiguous = oracle.render_f32(cam, projected, lists, background)
    actual = run.image_host()
    clear = ambiguous == 0
    assert clear.mean() > 0.99, f"ambiguous pixels: {(~clear).sum()}"
    error = np.abs(actual - expected)[clear]
    tolerance = 1e-5 * np.abs(expected[clear]) + 1e-6
    assert np.all(error <= tolerance), f"max error {error.max()}"


def _reference_image(oracle, params, cam, background, degree):
    return oracle.render_reference(params, cam, background, degree)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
@pytest.mark.parametrize("size", [(9, 7), (19, 17), (33, 31)])
def test_reference_case_sh_degrees_and_partial_tiles(gsn, context, oracle, scenes, degree, size):
    """renderer_test.cu:9-29."""
    params = scenes.make_gaussians(7)
    cam = reference_camera(gsn, *size)
    background = (0.05, 0.1, 0.15)
    run = GpuRun(gsn, context, params, cam, background, degree)
    expected = _reference_image(oracle, params, cam, background, degree)
    actual = run.image_host().astype(np.float64)
    assert np.all(np.isfinite(actual))
    assert np.abs(actual - expected).max() < GOLDEN["render_linf"]


def test_reference_case_empty_and_culled(gsn, context, scenes):
    """renderer_test.cu:31-46: culled and empty scenes give the exact background."""
    cam = reference_camera(gsn, 19, 17)
    background = (0.2, 0.4, 0.6)
    params = scenes.make_gaussians(2)
    params[2, 0] = 0.1
    params[2, 1] = 101.0
    for count in (2, 0):
        run = GpuRun(gsn, context, params[:, :count], cam, background, 0)
        image = run.image_host()
        np.testing.assert_array_equal(image, np.broadcast_to(np.float32(background), image.shape))


def test_reference_case_depth_order_and_termination(gsn, context, oracle, scenes):
    """renderer_test.cu:48-70."""
    params = scenes.make_gaussians(8)
    for i in range(8):
        params[0:3, i] = [0.0, 0.0, 3.0 + i * 0.1]
        params[3:6, i] = 0.0
        params[9, i] = np.log(np.float32(49.0))
        params[10, i] = (np.float32(0.2) + np.float32(i) * np.float32(0.08) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    params = params[:, ::-1].copy()
    cam = reference_camera(gsn, 9, 7)
    background = (0.1, 0.2, 0.3)
    run = GpuRun(gsn, context, params, cam, background, 0)
    expected = _reference_image(oracle, params, cam, background, 0)
    assert np.abs(run.image_host() - expected).max() < GOLDEN["render_linf"]


def test_reference_case_lower_colour_clamp(gsn, context, scenes):
    """renderer_test.cu:72-88: clamp below at 0 only; the centre pixel reads 0 and 1.35."""
    params = scenes.make_gaussians(1)
    params[0:3, 0] = [0.0, 0.0, 3.0]
    params[9, 0] = np.log(np.float32(9.0))
    params[10, 0] = (np.float32(-0.2) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    params[26, 0] = (np.float32(1.5) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    cam = reference_camera(gsn, 9, 7)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), 0)
    pixel = run.image_host()[3, 4]
    clamp = GOLDEN["color_clamp"]
    assert pixel[0] == clamp["red"]
    assert abs(pixel[1] - clamp["green"]) < clamp["tolerance"]


@pytest.mark.parametrize("count", [257, 513])
def test_reference_case_multiple_batches(gsn, context, oracle, scenes, count):
    """renderer_test.cu:89-114: tile lists longer than one batch."""
    params = scenes.make_gaussians(count)
    for i in range(count):
        params[0:3, i] = [0.1, -0.1, 3.0 + 0.001 * i]
        params[3:6, i] = np.log(np.float32(0.8))
        params[9, i] = -4.4
        params[10, i] += np.float32(0.1) * np.sin(np.float32(i))
    params = params[:, ::-1].copy()
    cam = reference_camera(gsn, 9, 7)
    background = (0.1, 0.2, 0.3)
    run = GpuRun(gsn, context, params, cam, background, 0)
    expected = _reference_image(oracle, params, cam, background, 0)
    assert np.abs(run.image_host() - expected).max() < GOLDEN["render_linf"]


def test_image_metrics_match_oracle_loss(gsn, context, oracle, scenes):
    rng = np.random.default_rng(5)
    image = rng.uniform(0, 1, size=(70, 45, 3)).astype(np.float32)
    target = rng.integers(0, 256, size=(70, 45, 3), dtype=np.uint8)
    metrics = gsn.image_metrics(context, 45, 70, gsn.DeviceBuffer.from_numpy(image), gsn.DeviceBuffer.from_numpy(target))
    expected = oracle.loss_f64(image.astype(np.float64), target)
    assert abs(metrics["loss"] - expected) < 1e-5 * max(1.0, abs(expected))
    mse = np.mean((np.clip(image, 0, 1).astype(np.float64) - target / 255.0) ** 2)
    assert abs(metrics["mse"] - mse) < 1e-6


def test_stage_timing_mask_times_only_selected_stages(gsn, scenes):
    """gsn_context_set_timing_mask records event pairs only around the selected stages; with
    every stage selected, each stage of a render records one launch."""
    context = gsn.Context()
    params = scenes.config0_params(seed=5, count=2000, sh_degree=0)
    cam = config0_camera(gsn)
    try:
        context.set_timing(True, stages=["raster_forward"])
        GpuRun(gsn, context, params, cam)
        times = context.stage_times()
        assert times["raster_forward"][1] == 1 and times["raster_forward"][0] > 0.0
        assert all(launches == 0 for name, (_, launches) in times.items() if name != "raster_forward")
        context.set_timing(True)
        GpuRun(gsn, context, params, cam)
        times = context.stage_times()
        for name in ("project", "scan", "depth_sort", "emit", "tile_sort", "ranges", "raster_forward"):
            assert times[name][1] >= 1, name
        context.set_timing(False)
        GpuRun(gsn, context, params, cam)
        assert all(launches == 0 for _, launches in context.stage_times().values())
    finally:
        context.close()
