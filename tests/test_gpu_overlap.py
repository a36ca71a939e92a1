"""The backward's preparation (packed splats and the two schedules) can be queued on a second
stream by an armed tile render and reused by the gradient call of the same projection. These
tests require results bit for bit equal to the one-stream path: over training iterations that
cross a density step, when the gradient call names another projection than the armed render's,
and when a projection runs between the armed render and the gradient call.
"""
import numpy as np
import pytest

from helpers import config0_camera

pytestmark = pytest.mark.gpu


def _bits(buffer):
    return buffer.numpy().view(np.uint32)


def test_training_with_and_without_overlap_is_identical(gsn):
    import gs_native_amd.trainer as trainer_module
    import gs_native_amd.workloads as workloads
    context = gsn.Context()
    params, cams, targets, extent, _ = workloads.build_workload(gsn, context, 20_000, 320, 240, 8)
    results = []
    for overlap in (False, True):
        before = context.prefetch_reused()
        trainer = trainer_module.Trainer(context, params, cams, targets, extent, completed=580)
        trainer.overlap_backward = overlap
        for _ in range(40):  # crosses the density step after iteration 600
            trainer.iteration()
        context.synchronize()
        assert context.prefetch_reused() - before == (40 if overlap else 0)
        results.append((trainer.count, trainer.params.numpy().copy(), trainer.moments.numpy()))
    print(f"count after the density step: {results[0][0]}", flush=True)
    assert results[0][0] == results[1][0]
    np.testing.assert_array_equal(results[0][1].view(np.uint32), results[1][1].view(np.uint32))
    for a, b in zip(results[0][2], results[1][2]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def _gradients(gsn, context, projection, params, image, image_grad):
    n = projection.count
    out = [gsn.DeviceBuffer.zeros(shape, np.float32) for shape in ((58, n), (2, n), (2, n), (11, n))]
    gsn.gaussian_gradients(context, projection, params, 3, image, image_grad, out[0], out[1], out[2], out[3])
    context.synchronize()
    return [_bits(b) for b in out]


def test_prefetch_of_another_projection_is_not_used(gsn, scenes):
    context = gsn.Context()
    n = 3000
    host = scenes.config0_params(seed=91, count=n, sh_degree=3)
    host[3:6] += np.float32(1.0)
    params = gsn.DeviceBuffer.from_numpy(host)
    cam_a = config0_camera(gsn, 192, 160, distance=3.0)
    cam_b = config0_camera(gsn, 192, 160, distance=2.5)
    rng = np.random.default_rng(5)
    image_grad = gsn.DeviceBuffer.from_numpy(rng.normal(0, 1e-3, (160, 192, 3)).astype(np.float32))
    image_a = gsn.DeviceBuffer((160, 192, 3), np.float32)
    image_b = gsn.DeviceBuffer((160, 192, 3), np.float32)
    projection_a = gsn.render(context, params, n, cam_a, (0, 0, 0), image_a, 3)
    reference_a = _gradients(gsn, context, projection_a, params, image_a, image_grad)
    projection_b = gsn.render(context, params, n, cam_b, (0, 0, 0), image_b, 3)
    reference_b = _gradients(gsn, context, projection_b, params, image_b, image_grad)
    assert projection_a.record_count != projection_b.record_count

    assert context.prefetch_reused() == 0

    # Armed render of B, then the gradients of A: the queued work is B's, so it is not used.
    context.prefetch_backward()
    projection_b = gsn.render(context, params, n, cam_b, (0, 0, 0), image_b, 3, projection_b)
    got = _gradients(gsn, context, projection_a, params, image_a, image_grad)
    for g, r in zip(got, reference_a):
        np.testing.assert_array_equal(g, r)
    assert context.prefetch_reused() == 0

    # Armed render of B, then the gradients of B: the queued work is used.
    context.prefetch_backward()
    projection_b = gsn.render(context, params, n, cam_b, (0, 0, 0), image_b, 3, projection_b)
    got = _gradients(gsn, context, projection_b, params, image_b, image_grad)
    for g, r in zip(got, reference_b):
        np.testing.assert_array_equal(g, r)
    assert context.prefetch_reused() == 1

    # Armed render of B, a projection of A in between, then the gradients of B: the queued work
    # belongs to an older projection and is redone.
    context.prefetch_backward()
    projection_b = gsn.render(context, params, n, cam_b, (0, 0, 0), image_b, 3, projection_b)
    projection_a = gsn.render(context, params, n, cam_a, (0, 0, 0), image_a, 3, projection_a)
    got = _gradients(gsn, context, projection_b, params, image_b, image_grad)
    for g, r in zip(got, reference_b):
        np.testing.assert_array_equal(g, r)
    assert context.prefetch_reused() == 1
