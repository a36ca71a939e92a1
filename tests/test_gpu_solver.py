"""Solver parity on the GPU: fused L1+SSIM gradient, recompute backward, chain rule, Adam and
history remap against the oracle, plus the reference's own solver cases (test/solver_test.cu).

Tolerance rule for gradients (sums over many pixels cancel): |gpu - oracle| <= 1e-5 * sum|terms|
+ 1e-9 * max|row|, where sum|terms| is the oracle's sum of the per-pixel terms' error-bound
magnitudes, i.e. 1e-5 relative to what an FP32 evaluation of the same algorithm can resolve:
  - every term is weighted by 1 + kappa, kappa = sum alpha_i / (1 - alpha_i) over the splats in
    front: the relative condition number of T = prod(1 - alpha_i) (1 - alpha loses bits as alpha
    nears the 0.99 cap);
  - for the geometry fields, |g| (|prefix| + |T c| + |C|) / (1 - alpha), the summands of dL/dalpha
    before they cancel, stands in for |dL/dalpha|: the replay from the final colour
    (solver.cu:389-396) subtracts nearly equal FP32 colours behind the last splats of a pixel.
The reference's own FP32 kernel has the same error behaviour.
"""
import json
import pathlib

import numpy as np
import pytest

from helpers import GpuRun, config0_camera, reference_camera, solver_step

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
SCREEN_TOLERANCE = 1e-5


@pytest.mark.parametrize("size", [(33, 65), (64, 48), (300, 200)])
def test_image_gradients_match_fp64(gsn, context, oracle, size):
    width, height = size
    rng = np.random.default_rng(width)
    image = rng.uniform(-0.1, 1.1, size=(height, width, 3)).astype(np.float32)
    target = rng.integers(0, 256, size=(height, width, 3), dtype=np.uint8)
    cam = reference_camera(gsn, width, height)
    out = gsn.DeviceBuffer((height, width, 3), np.float32)
    gsn.image_gradients(context, cam, gsn.DeviceBuffer.from_numpy(image), gsn.DeviceBuffer.from_numpy(target), out)
    context.synchronize()
    expected = oracle.image_gradients_f64(image.astype(np.float64), target)
    actual = out.numpy()
    scale = 1.0 / (3.0 * width * height)
    assert np.abs(actual - expected).max() <= 1e-5 * scale


def _scene_and_target(gsn, oracle, scenes, seed, count, width, sh_degree):
    params = scenes.config0_params(seed=seed, count=count, sh_degree=3)
    params[3:6] += np.float32(1.0)  # larger splats at this resolution
    cam = config0_camera(gsn, width, width)
    truth = scenes.perturb(params, seed + 1, 0.01)
    image = oracle.render_reference(truth, cam, (0.0, 0.0, 0.0), sh_degree)
    target = np.clip(np.rint(image * 255.0), 0, 255).astype(np.uint8)
    return params, cam, target


@pytest.mark.parametrize("sh_degree", [0, 3])
def test_screen_gradients_match_fp64_backward(gsn, context, oracle, scenes, sh_degree):
    params, cam, target = _scene_and_target(gsn, oracle, scenes, 31 + sh_degree, 2000, 128, sh_degree)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), sh_degree)
    got = run.gradients(target)
    f32 = oracle.project_f32(params, cam, sh_degree)
    # The FP64 replay runs on the FP32 projection (bit-identical to the device's) and on the
    # device's dL/dimage, so that the comparison measures the rasteriser backward alone: the
    # projected means carry an absolute FP32 error of ~|mean| * 6e-8 px, which moves alpha by |q|
    # times that for small splats, and the SSIM gradient's local variances cancel in flat regions
    # (both inputs are checked against FP64 by their own tests).
    ref = oracle.raster_gradients_f64(cam, f32, (0.0, 0.0, 0.0), target, image_grad=got["image_grad"])
    # Gaussians that contribute to a pixel where some threshold decision (alpha vs 1/255, the
    # 0.99 cap, T vs 1e-4) lies within 1e-4 relative of its threshold may legitimately differ
    # between FP32 and FP64; in this dense scene every such pixel has ~40 contributors.
    keep = ~ref["tainted"]
    assert keep.mean() > 0.6
    screen = got["screen"]
    for k in range(11):
        error = np.abs(screen[k] - ref["screen"][k])[keep]
        bound = SCREEN_TOLERANCE * ref["magnitude"][k][keep] + 1e-9 * np.abs(ref["screen"][k]).max()
        worst = np.argmax(error - bound)
        assert np.all(error <= bound), f"field {k}: error {error[worst]} bound {bound[worst]}"
    np.testing.assert_array_equal(got["statistics"][1], (f32["record_counts"][:-1] > 0).astype(np.float32))


@pytest.mark.parametrize("sh_degree", [0, 1, 2, 3])
def test_chain_rule_matches_fp64_map(gsn, context, oracle, scenes, sh_degree):
    """The parameter gradients equal the FP64 chain rule applied to the GPU's own screen-space
    gradients (isolates the map from the backward)."""
    params, cam, target = _scene_and_target(gsn, oracle, scenes, 41 + sh_degree, 2000, 128, sh_degree)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), sh_degree)
    got = run.gradients(target)
    observed = (np.diff(run.lists()["record_offsets"]) > 0).astype(np.uint8)
    ref = oracle.map_f64(params, cam, sh_degree, got["screen"].astype(np.float64), observed)
    for row in range(58):
        expected = ref["gradients"][row]
        error = np.abs(got["gradients"][row] - expected)
        bound = 1e-5 * np.abs(expected) + 1e-5 * np.abs(expected).max() + 1e-30
        assert np.all(error <= bound), f"row {row}: max error {error.max()} of {np.abs(expected).max()}"
    np.testing.assert_allclose(got["statistics"][0], ref["statistics"][0], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(got["absolute_statistics"][0], ref["absolute_statistics"][0], rtol=1e-5, atol=1e-12)


def test_adam_bit_exact(gsn, context, oracle):
    rng = np.random.default_rng(3)
    n = 5000
    gradients = rng.normal(0, 1e-3, size=(58, n)).astype(np.float32)
    gradients[:, ::7] = 0.0
    params = rng.normal(size=(58, n)).astype(np.float32)
    first = rng.normal(0, 1e-3, size=(58, n)).astype(np.float32)
    second = rng.uniform(0, 1e-6, size=(58, n)).astype(np.float32)
    p_dev = gsn.DeviceBuffer.from_numpy(params)
    moments = gsn.Moments.from_numpy(first, second)
    gsn.adam_update(context, n, gsn.DeviceBuffer.from_numpy(gradients), p_dev, moments, 1.6e-4, 7)
    context.synchronize()
    c1, c2 = gsn.adam_corrections(7)
    ep, em, ev = oracle.adam_f32(gradients, params, first, second, 1.6e-4, c1, c2)
    got_first, got_second = moments.numpy()
    np.testing.assert_array_equal(p_dev.numpy().view(np.uint32), ep.view(np.uint32))
    np.testing.assert_array_equal(got_first.view(np.uint32), em.view(np.uint32))
    np.testing.assert_array_equal(got_second.view(np.uint32), ev.view(np.uint32))


@pytest.mark.parametrize("sh_degree", [0, 3])
def test_fused_step_equals_gradients_then_adam(gsn, context, scenes, oracle, sh_degree):
    params, cam, target = _scene_and_target(gsn, oracle, scenes, 51, 3000, 128, sh_degree)
    rng = np.random.default_rng(9)
    first = rng.normal(0, 1e-4, size=params.shape).astype(np.float32)
    second = rng.uniform(0, 1e-8, size=params.shape).astype(np.float32)
    fused = solver_step(gsn, context, params, cam, target, sh_degree, first=first, second=second, step=3)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), sh_degree)
    got = run.gradients(target, screen=False)
    c1, c2 = gsn.adam_corrections(3)
    ep, em, ev = oracle.adam_f32(got["gradients"], params, first, second, 1e-4, c1, c2)
    np.testing.assert_array_equal(fused["params"].view(np.uint32), ep.view(np.uint32))
    np.testing.assert_array_equal(fused["first"].view(np.uint32), em.view(np.uint32))
    np.testing.assert_array_equal(fused["second"].view(np.uint32), ev.view(np.uint32))


@pytest.mark.parametrize("count", [33, 257, 20_011])
def test_fused_step_at_counts_off_the_line_grid(gsn, context, scenes, oracle, count):
    """A count that is not a multiple of 32 puts the parameter and moment rows off the 128-byte
    line grid, where the fused pass uses ordinary instead of non-temporal accesses; 33 and 257 also
    leave a partial last workgroup. The fused pass must equal the unfused gradients followed by
    Adam, and its statistics the unfused pass's, bit for bit."""
    params, cam, target = _scene_and_target(gsn, oracle, scenes, 61, count, 128, 3)
    rng = np.random.default_rng(count)
    first = rng.normal(0, 1e-4, size=params.shape).astype(np.float32)
    second = rng.uniform(0, 1e-8, size=params.shape).astype(np.float32)
    fused = solver_step(gsn, context, params, cam, target, 3, first=first, second=second, step=4)
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), 3)
    got = run.gradients(target, screen=False)
    c1, c2 = gsn.adam_corrections(4)
    ep, em, ev = oracle.adam_f32(got["gradients"], params, first, second, 1e-4, c1, c2)
    assert np.abs(got["gradients"]).max() > 0
    np.testing.assert_array_equal(fused["params"].view(np.uint32), ep.view(np.uint32))
    np.testing.assert_array_equal(fused["first"].view(np.uint32), em.view(np.uint32))
    np.testing.assert_array_equal(fused["second"].view(np.uint32), ev.view(np.uint32))
    np.testing.assert_array_equal(fused["statistics"].view(np.uint32), got["statistics"].view(np.uint32))
    np.testing.assert_array_equal(fused["absolute_statistics"].view(np.uint32),
                                  got["absolute_statistics"].view(np.uint32))


def test_remap_history_bit_exact(gsn, context, oracle):
    rng = np.random.default_rng(4)
    previous, count = 700, 1000
    first = rng.normal(size=(58, previous)).astype(np.float32)
    second = rng.normal(size=(58, previous)).astype(np.float32)
    parents = rng.integers(-3, previous + 3, size=count).astype(np.int32)
    out = gsn.Moments(count)
    # Keep every input alive until the stream has finished with it.
    parents_dev = gsn.DeviceBuffer.from_numpy(parents)
    previous_moments = gsn.Moments.from_numpy(first, second)
    gsn.call("gsn_remap_history", context.handle, parents_dev.ptr, previous, count, previous_moments.ptr, out.ptr)
    context.synchronize()
    ef, es = oracle.remap_f32(parents, first, second, count)
    got_first, got_second = out.numpy()
    np.testing.assert_array_equal(got_first, ef)
    np.testing.assert_array_equal(got_second, es)


# ---- reference solver cases (test/solver_test.cu) -------------------------------------------

def _differentiable_scene(gsn, scenes, width=9, height=7):
    rotation = scenes.differentiable_scene_rotation()
    cam = reference_camera(gsn, width, height, rotation)
    params = scenes.make_gaussians(1)
    params[0:3, 0] = rotation.T @ np.array([0.2, -0.1, 3.0], np.float32)
    params[3:6, 0] = np.log(np.array([0.6, 0.45, 0.35], np.float32))
    return params, cam


def _loss(oracle, params, cam, target, degree, background=(0.0, 0.0, 0.0)):
    return oracle.loss_f64(oracle.render_reference(params, cam, background, degree), target)


def _fd(oracle, params, cam, target, degree, index, field, background=(0.0, 0.0, 0.0), h=0.0005):
    high, low = params.copy(), params.copy()
    high[field, index] += np.float32(h)
    low[field, index] -= np.float32(h)
    step = float(high[field, index]) - float(low[field, index])
    return (_loss(oracle, high, cam, target, degree, background) - _loss(oracle, low, cam, target, degree, background)) / step


def _active(field, degree):
    return field < 10 or (field - 10) % 16 < (degree + 1) ** 2


def _constant_target(width, height, values=(180, 190, 200)):
    target = np.zeros((height, width, 3), np.uint8)
    target[:] = values
    return target


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_reference_case_update_directions_match_fd(gsn, context, oracle, scenes, degree):
    """solver_test.cu:58-95: every active field moves against the FD loss derivative."""
    params, cam = _differentiable_scene(gsn, scenes)
    target = _constant_target(9, 7)
    result = solver_step(gsn, context, params, cam, target, degree)
    after = result["params"]
    assert np.all(np.isfinite(after))
    verified = 0
    for field in range(58):
        if not _active(field, degree):
            assert after[field, 0] == params[field, 0]
            continue
        derivative = _fd(oracle, params, cam, target, degree, 0, field)
        assert abs(derivative) > 1e-9, f"field {field}"
        assert derivative * (after[field, 0] - params[field, 0]) < 0, f"field {field}"
        verified += 1
    assert verified == 10 + 3 * (degree + 1) ** 2


@pytest.mark.parametrize("degree", [0, 3])
def test_reference_case_overlapping_with_background(gsn, context, oracle, scenes, degree):
    """solver_test.cu:97-154."""
    base, cam = _differentiable_scene(gsn, scenes)
    rotation = cam.rotation()
    params = scenes.make_gaussians(3)
    for i in range(3):
        params[0:3, i] = rotation.T @ np.array([0.2 - 0.12 * i, -0.1 + 0.09 * i, 3.0 + 0.4 * i], np.float32)
        params[3:6, i] = base[3:6, 0]
        params[9, i] = -0.4 + 0.3 * i
        for c in range(3):
            params[10 + 16 * c, i] = (0.25 + 0.15 * ((i + c) % 3) - 0.5) / scenes.SH_C0
    params = params[:, ::-1].copy()
    background = (0.12, 0.31, 0.57)
    target = _constant_target(9, 7)
    result = solver_step(gsn, context, params, cam, target, degree, background)
    after = result["params"]
    for i in range(3):
        for field in range(58):
            if not _active(field, degree):
                assert after[field, i] == params[field, i]
                continue
            derivative = _fd(oracle, params, cam, target, degree, i, field, background)
            assert abs(derivative) > 1e-9
            assert derivative * (after[field, i] - params[field, i]) < 0, f"gaussian {i} field {field}"


def _statistic_fd(oracle, params, cam, target, degree):
    derivatives = []
    for axis in range(2):
        high = oracle.as_camera(cam)
        low = oracle.as_camera(cam)
        if axis == 0:
            high.cx += 0.001
            low.cx -= 0.001
            difference = float(np.float32(high.cx)) - float(np.float32(low.cx))
        else:
            high.cy += 0.001
            low.cy -= 0.001
            difference = float(np.float32(high.cy)) - float(np.float32(low.cy))
        derivatives.append((_loss(oracle, params, high, target, degree) - _loss(oracle, params, low, target, degree)) / difference)
    return float(np.hypot(derivatives[0] * cam.width * 0.5, derivatives[1] * cam.height * 0.5))


def test_reference_case_projection_statistic_fd(gsn, context, oracle, scenes):
    """solver_test.cu:156-178."""
    params, cam = _differentiable_scene(gsn, scenes)
    target = _constant_target(9, 7)
    result = solver_step(gsn, context, params, cam, target, 3)
    expected = _statistic_fd(oracle, params, cam, target, 3)
    tol = GOLDEN["statistic_fd_tolerance"]
    assert abs(result["statistics"][0, 0] - expected) < tol["absolute"] + expected * tol["relative"]
    assert result["statistics"][1, 0] == 1.0


def test_reference_case_statistic_across_ssim_blocks(gsn, context, oracle, scenes):
    """solver_test.cu:276-307: 33x65 image, large splat, patterned target."""
    params, cam = _differentiable_scene(gsn, scenes, 33, 65)
    params[3:6, 0] = np.log(np.array([3.0, 2.5, 1.8], np.float32))
    y, x = np.mgrid[0:65, 0:33]
    target = np.stack([120 + (3 * x + y) % 100, 100 + (x + 2 * y) % 120, 140 + (2 * x + 3 * y) % 100], -1).astype(np.uint8)
    result = solver_step(gsn, context, params, cam, target, 3)
    expected = _statistic_fd(oracle, params, cam, target, 3)
    assert expected > 1e-6
    tol = GOLDEN["statistic_fd_tolerance"]
    assert abs(result["statistics"][0, 0] - expected) < tol["absolute"] + expected * tol["relative"]


def test_reference_case_parent_permutation(gsn, context, scenes):
    """solver_test.cu:180-202: a permuted remap keeps subsequent updates permuted."""
    params = scenes.make_gaussians(2)
    cam = reference_camera(gsn, 19, 17)
    target_a = np.full((17, 19, 3), 180, np.uint8)
    target_b = np.full((17, 19, 3), 20, np.uint8)
    first = solver_step(gsn, context, params, cam, target_a, 1)
    second = solver_step(gsn, context, params, cam, target_a, 1)
    swapped = second["params"][:, ::-1].copy()
    parents = gsn.DeviceBuffer.from_numpy(np.array([1, 0], np.int32))
    remapped = gsn.Moments(2)
    previous = gsn.Moments.from_numpy(second["first"], second["second"])
    gsn.call("gsn_remap_history", context.handle, parents.ptr, 2, 2, previous.ptr, remapped.ptr)
    context.synchronize()
    m, v = remapped.numpy()
    a = solver_step(gsn, context, first["params"], cam, target_b, 1, first=first["first"], second=first["second"], step=2)
    b = solver_step(gsn, context, swapped, cam, target_b, 1, first=m, second=v, step=2)
    np.testing.assert_allclose(a["params"], b["params"][:, ::-1], atol=2e-6)


def test_reference_case_zero_history_first_step(gsn, context, scenes):
    """solver_test.cu:204-236: after an opacity-history reset (or a new parent) the opacity moves
    by exactly lr * (0.1 c1) / sqrt(0.001 c2) at step 2."""
    params, cam = _differentiable_scene(gsn, scenes)
    target = np.full((7, 9, 3), 200, np.uint8)
    first = solver_step(gsn, context, params, cam, target, 0)
    values = first["params"].copy()
    values[9, 0] = np.log(np.float32(0.01) / np.float32(0.99))
    m, v = first["first"].copy(), first["second"].copy()
    m[9] = 0.0
    v[9] = 0.0
    second = solver_step(gsn, context, values, cam, target, 0, first=m, second=v, step=2)
    factor = (0.1 / (1 - 0.9 ** 2)) / np.sqrt(0.001 / (1 - 0.999 ** 2))
    expected = GOLDEN["adam_reset_step"]["learning_rate"] * factor
    assert abs(abs(second["params"][9, 0] - values[9, 0]) - expected) < GOLDEN["adam_reset_step"]["tolerance"]


def test_reference_case_gradients_cross_batch_boundary(gsn, context, oracle, scenes):
    """solver_test.cu:237-274: 257 overlapping Gaussians, update signs at batch edges."""
    count = 257
    params = scenes.make_gaussians(count)
    for i in range(count):
        params[0:3, i] = [0.1, -0.1, 3.0 + 0.001 * i]
        params[3:6, i] = np.log(np.float32(0.8))
        params[9, i] = -4.4
        params[10, i] += np.float32(0.1) * np.sin(np.float32(i))
    params = params[:, ::-1].copy()
    cam = reference_camera(gsn, 9, 7)
    target = _constant_target(9, 7)
    result = solver_step(gsn, context, params, cam, target, 0)
    after = result["params"]
    assert np.all(np.isfinite(after))
    for i in (0, 1, 128, 255, 256):
        for field in (9, 10, 26, 42):
            derivative = _fd(oracle, params, cam, target, 0, i, field)
            assert abs(derivative) > 1e-9
            assert derivative * (after[field, i] - params[field, i]) < 0, f"gaussian {i} field {field}"


def test_reference_case_absolute_statistic(gsn, context, oracle, scenes):
    """solver_test.cu:309-356: AbsGS statistic against pixel-wise finite differences."""
    params = scenes.make_gaussians(1)
    params[0:3, 0] = [0.0, 0.0, 3.0]
    params[3:6, 0] = np.log(np.float32(0.5))
    params[10:58, 0] = 0.0
    cam = reference_camera(gsn, 9, 7)
    target = _constant_target(9, 7)
    reference = oracle.render_reference(params, cam, (0, 0, 0), 0)
    colour_derivatives = []
    for axis in range(2):
        high, low = oracle.as_camera(cam), oracle.as_camera(cam)
        if axis == 0:
            high.cx += 0.001
            low.cx -= 0.001
            step = float(np.float32(high.cx)) - float(np.float32(low.cx))
        else:
            high.cy += 0.001
            low.cy -= 0.001
            step = float(np.float32(high.cy)) - float(np.float32(low.cy))
        colour_derivatives.append((oracle.render_reference(params, high, (0, 0, 0), 0)
                                   - oracle.render_reference(params, low, (0, 0, 0), 0)) / step)
    sums = [0.0, 0.0]
    for y in range(7):
        for x in range(9):
            contribution = [0.0, 0.0]
            for c in range(3):
                high, low = reference.copy(), reference.copy()
                high[y, x, c] += 1e-4
                low[y, x, c] -= 1e-4
                g = (oracle.loss_f64(high, target) - oracle.loss_f64(low, target)) / 2e-4
                for axis in range(2):
                    contribution[axis] += g * colour_derivatives[axis][y, x, c]
            for axis in range(2):
                sums[axis] += abs(contribution[axis])
    expected = float(np.hypot(sums[0] * 9 * 0.5, sums[1] * 7 * 0.5))
    result = solver_step(gsn, context, params, cam, target, 0)
    assert expected > 1e-5
    tol = GOLDEN["statistic_fd_tolerance"]
    assert abs(result["absolute_statistics"][0, 0] - expected) < tol["absolute"] + expected * tol["relative"]
    assert result["absolute_statistics"][0, 0] > result["statistics"][0, 0] + 1e-5
