"""The 500,000-Gaussian GPU configuration: a radius-2 shell plus a ground plane, SH degree 3
(higher bands N(0, 0.05)), 1248x832, 150 cameras on a radius-4 hemisphere, trained from a
seed-perturbed copy with the reference schedule (src/core/gaussian_splatting.cpp:137-189).

- Iterations 551-650 cross the density step at iteration 600. The state going into that step
  (parameters, accumulated gradient sums, observation counts, Adam moments) is read back, and
  the step's outputs are compared with the oracle bit for bit: the decisions' offsets, the new
  parameters (keep / clone / split with the split jitter), the history parents and the remapped
  Adam moments.
- Two runs of iterations 591-610 from the same state agree bit for bit (density step included).
- One view of the trained state keeps the tile-list invariants at full size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNT = 500_000
WIDTH, HEIGHT = 1248, 832
CAMERAS = 150


@pytest.fixture(scope="module")
def config1(gsn, context):
    import gs_native_amd.workloads as workloads
    params, cams, targets, extent, _ = workloads.build_workload(
        gsn, context, COUNT, WIDTH, HEIGHT, CAMERAS, seed=31, scene="shell", camera_radius=4.0, hemisphere=True)
    return params, cams, targets, extent


def _trainer(gsn, context, config1, completed):
    import gs_native_amd.trainer as trainer_module
    params, cams, targets, extent = config1
    return trainer_module.Trainer(context, params, cams, targets, extent, completed=completed)


def test_density_step_bit_exact_at_config1(gsn, context, oracle, config1):
    trainer = _trainer(gsn, context, config1, completed=550)
    captured = {}
    control = trainer.control_density

    def capturing_control(*args, **kwargs):
        n = trainer.count
        captured["before"] = {
            "count": n, "completed": trainer.completed, "counter": trainer.density_counter,
            "params": trainer.params.numpy()[:, :n].copy(), "sums": trainer.sums.numpy()[:n].copy(),
            "absolute_sums": trainer.absolute_sums.numpy()[:n].copy(),
            "observations": trainer.observations.numpy()[:n].copy(),
            "moments": trainer.moments.numpy(),
        }
        changes = control(*args, **kwargs)
        m = trainer.count
        captured["after"] = {"count": m, "params": trainer.params.numpy()[:, :m].copy(),
                             "moments": trainer.moments.numpy()}
        return changes

    trainer.control_density = capturing_control
    for _ in range(100):
        trainer.iteration()
    context.synchronize()
    assert captured, "no density step ran in iterations 551-650"
    before, after = captured["before"], captured["after"]
    assert before["completed"] == 600
    n = before["count"]
    assert n == COUNT
    assert before["observations"].max() > 0
    extent = np.float32(trainer.scene_extent)
    split_threshold = float(np.float32(0.01) * extent)
    world = float(np.float32(0.1) * extent)
    decisions, offsets = oracle.density_count_f32(before["params"], before["sums"], before["absolute_sums"],
                                                  before["observations"], 0.0002, split_threshold, 0.005, world,
                                                  False, 0.0008)
    totals = offsets[n]
    expected_count = int(totals[0] + totals[1] + 2 * totals[3])
    print(f"density step at 600: {n} -> {after['count']} (keep {totals[0]}, clone {totals[1]}, split {totals[3]})",
          flush=True)
    assert after["count"] == expected_count
    assert totals[1] + totals[3] > 0, "the step should densify something"
    out, history, _, _ = oracle.density_write_f32(before["params"], decisions, offsets, 0, before["counter"])
    np.testing.assert_array_equal(after["params"].view(np.uint32), out.view(np.uint32))
    first, second = oracle.remap_f32(history, *before["moments"], expected_count)
    np.testing.assert_array_equal(after["moments"][0].view(np.uint32), first.view(np.uint32))
    np.testing.assert_array_equal(after["moments"][1].view(np.uint32), second.view(np.uint32))


def test_config1_training_is_deterministic(gsn, context, config1):
    results = []
    for _ in range(2):
        trainer = _trainer(gsn, context, config1, completed=590)
        for _ in range(20):
            trainer.iteration()
        context.synchronize()
        assert trainer.history and trainer.history[0][0] == 600
        results.append((trainer.count, trainer.params_host().copy(), trainer.moments.numpy()[0]))
        del trainer
    assert results[0][0] == results[1][0]
    np.testing.assert_array_equal(results[0][1].view(np.uint32), results[1][1].view(np.uint32))
    np.testing.assert_array_equal(results[0][2].view(np.uint32), results[1][2].view(np.uint32))


def test_config1_tile_lists_invariants(gsn, context, config1):
    params, cams, _, _ = config1
    device_params = gsn.DeviceBuffer.from_numpy(params)
    image = gsn.DeviceBuffer((HEIGHT, WIDTH, 3), np.float32)
    cam = cams[7]
    projection = gsn.render(context, device_params, COUNT, cam, (0.0, 0.0, 0.0), image, 3, keep_keys=True)
    context.synchronize()
    r = projection.record_count
    assert r > COUNT
    keys = projection.sorted_keys.numpy()[:r]
    ids = projection.tile_gaussian_ids.numpy()[:r]
    offsets = projection.tile_offsets.numpy()[: cam.tile_count + 1]
    counts = np.diff(projection.record_offsets.numpy())
    rects = projection.tile_rects.numpy()[:COUNT]
    assert np.all(keys[1:] >= keys[:-1])
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    np.testing.assert_array_equal(offsets, np.searchsorted(tiles, np.arange(cam.tile_count + 1), side="left"))
    visible = counts > 0
    areas = (rects[:, 2] - rects[:, 0]).astype(np.int64) * (rects[:, 3] - rects[:, 1])
    np.testing.assert_array_equal(counts[visible], areas[visible])
    tx, ty = tiles % cam.tile_count_x, tiles // cam.tile_count_x
    rect = rects[ids]
    assert np.all((tx >= rect[:, 0]) & (tx < rect[:, 2]) & (ty >= rect[:, 1]) & (ty < rect[:, 3]))
    same = keys[1:] == keys[:-1]
    assert np.all(ids[1:][same] > ids[:-1][same])
    assert np.all(np.isfinite(image.numpy()))
