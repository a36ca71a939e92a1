"""Pins the CPU oracle (no GPU): its transcendentals against libm, its image/loss math against an
independent scipy implementation and finite differences, its analytic gradients against FP64
finite differences of its own loss, its sort/compaction against brute force, its RNG against
published Philox4x32-10 known-answer vectors, and the reference tests' literal vectors."""
import json
import math
import pathlib

import numpy as np
import pytest
from scipy import ndimage

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ai = a.astype(np.float32).view(np.int32).astype(np.int64)
    bi = b.astype(np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def test_expf_within_two_ulp(oracle):
    xs = np.concatenate([np.linspace(-87, 88, 20001), np.linspace(-1, 1, 2001), [0.0, -0.0]]).astype(np.float32)
    got = np.array([oracle.expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64)).astype(np.float32)
    assert ulp_distance(got, want).max() <= 2


def test_logf_within_two_ulp(oracle):
    xs = np.concatenate([np.geomspace(1e-30, 1e30, 20001), np.linspace(0.5, 2.0, 4001)]).astype(np.float32)
    got = np.array([oracle.logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64)).astype(np.float32)
    assert ulp_distance(got, want).max() <= 2
    assert oracle.logf(1.0) == 0.0


def test_sincosf_accuracy(oracle):
    xs = np.concatenate([np.linspace(-20, 20, 20001), np.linspace(-1e-3, 1e-3, 101)]).astype(np.float32)
    for x in xs[::3]:
        s, c = oracle.sincosf(float(x))
        xd = float(x)
        assert abs(s - math.sin(xd)) <= 2.5 * np.spacing(np.float32(max(abs(math.sin(xd)), 1e-30))) + 1e-9
        assert abs(c - math.cos(xd)) <= 2.5 * np.spacing(np.float32(max(abs(math.cos(xd)), 1e-30))) + 1e-9


def test_ssim_window_matches_reference_constants(oracle):
    """The oracle's window (exp(-x^2/4.5) normalised) equals the reference's FP32 literals."""
    x = np.arange(11) - 5.0
    window = np.exp(-x * x / (2 * 1.5 * 1.5))
    window /= window.sum()
    np.testing.assert_allclose(window.astype(np.float32), GOLDEN["ssim_window"], rtol=1e-7)


def scipy_loss(image, target):
    x = np.arange(11) - 5.0
    w = np.exp(-x * x / 4.5)
    w /= w.sum()

    def filt(a):
        return ndimage.correlate1d(ndimage.correlate1d(a, w, axis=0, mode="constant"), w, axis=1, mode="constant")

    b = target / 255.0
    l1 = np.abs(image - b).mean()
    ssim = []
    for c in range(3):
        p, q = image[..., c], b[..., c]
        ma, mb = filt(p), filt(q)
        va, vb, cov = filt(p * p) - ma * ma, filt(q * q) - mb * mb, filt(p * q) - ma * mb
        ssim.append((2 * ma * mb + 1e-4) * (2 * cov + 9e-4) / ((ma * ma + mb * mb + 1e-4) * (va + vb + 9e-4)))
    return 0.8 * l1 + 0.2 * (1 - np.mean(ssim))


def test_loss_matches_independent_implementation(oracle):
    rng = np.random.default_rng(0)
    image = rng.uniform(0, 1, size=(23, 37, 3))
    target = rng.integers(0, 256, size=(23, 37, 3), dtype=np.uint8)
    assert abs(oracle.loss_f64(image, target) - scipy_loss(image, target)) < 1e-12
    same = target / 255.0
    assert abs(oracle.loss_f64(same, target)) < 1e-12


def test_image_gradients_match_fd(oracle):
    rng = np.random.default_rng(1)
    image = rng.uniform(0, 1, size=(13, 17, 3))
    target = rng.integers(0, 256, size=(13, 17, 3), dtype=np.uint8)
    grad = oracle.image_gradients_f64(image, target)
    h = 1e-6
    for (y, x, c) in [(0, 0, 0), (6, 8, 1), (12, 16, 2), (3, 15, 0), (11, 1, 1)]:
        high, low = image.copy(), image.copy()
        high[y, x, c] += h
        low[y, x, c] -= h
        fd = (oracle.loss_f64(high, target) - oracle.loss_f64(low, target)) / (2 * h)
        assert abs(fd - grad[y, x, c]) < 1e-6 * abs(grad[y, x, c]) + 1e-10


def _camera(oracle, width, height, rotation=None, translation=(0.0, 0.0, 0.0)):
    cam = oracle.Camera()
    cam.fx, cam.fy, cam.cx, cam.cy = 20.0, 21.0, 0.5 * (width - 1), 0.5 * (height - 1)
    rotation = np.eye(3) if rotation is None else rotation
    for k, v in enumerate(np.asarray(rotation, np.float32).reshape(9)):
        cam.R_c_w[k] = float(v)
    for k, v in enumerate(translation):
        cam.t_c_w[k] = float(v)
    cam.near_depth, cam.far_depth = 0.2, 100.0
    cam.width, cam.height = width, height
    cam.tile_count_x, cam.tile_count_y = (width + 15) // 16, (height + 15) // 16
    return cam


def _scenes():
    import importlib.util
    import sys
    spec = importlib.util.spec_from_file_location(
        "gsn_scenes", pathlib.Path(__file__).resolve().parents[1] / "gs-native_amd" / "scenes.py")
    module = importlib.util.module_from_spec(spec)
    sys.modules["gsn_scenes"] = module
    spec.loader.exec_module(module)
    return module


def test_single_gaussian_closed_form(oracle):
    scenes = _scenes()
    params = scenes.make_gaussians(1)
    params[0:3, 0] = [0.0, 0.0, 3.0]
    params[6:9, 0] = 0.0
    params[3:6, 0] = np.log(np.float32(0.3))
    cam = _camera(oracle, 9, 7)
    image = oracle.render_reference(params, cam, (0.1, 0.2, 0.3), 0)
    # Closed form: isotropic covariance s^2 I, J = diag(fx/z, fy/z) at the centre.
    s2 = np.exp(2.0 * float(params[3, 0]))
    z = 3.0
    a, c = (20.0 / z) ** 2 * s2 + 0.3, (21.0 / z) ** 2 * s2 + 0.3
    opacity = 1 / (1 + np.exp(1.0))
    colour = 0.5 + 0.28209479177387814 * params[10:58:16, 0].astype(np.float64)
    for (y, x) in [(3, 4), (0, 0), (6, 8)]:
        dx, dy = 4.0 - x, 3.0 - y
        # Off-axis pixels see the clamped Jacobian only through J's third column (zero at the centre).
        alpha = min(0.99, opacity * np.exp(-0.5 * (dx * dx / a + dy * dy / c)))
        expected = colour * alpha + (1 - alpha) * np.array([0.1, 0.2, 0.3], np.float32).astype(np.float64)
        np.testing.assert_allclose(image[y, x], expected, rtol=1e-12, atol=1e-12)


def _differentiable(scenes, oracle, width=9, height=7):
    rotation = scenes.differentiable_scene_rotation()
    cam = _camera(oracle, width, height, rotation)
    params = scenes.make_gaussians(1)
    params[0:3, 0] = rotation.T @ np.array([0.2, -0.1, 3.0], np.float32)
    params[3:6, 0] = np.log(np.array([0.6, 0.45, 0.35], np.float32))
    return params, cam


@pytest.mark.parametrize("degree", [0, 3])
def test_analytic_gradients_match_fp64_finite_differences(oracle, degree):
    """FP64 chain (image gradient -> replay backward -> chain rule) against central differences
    of loss(render(params)); pins the backward and the map, including rotation and SH terms."""
    scenes = _scenes()
    params, cam = _differentiable(scenes, oracle, 19, 17)
    params[10:58, 0] += np.linspace(-0.05, 0.05, 48).astype(np.float32)
    target = np.zeros((17, 19, 3), np.uint8)
    target[:] = (180, 190, 200)
    result = oracle.gradients_f64(params, cam, degree, (0.1, 0.1, 0.1), target)
    assert not result["tainted"].any()
    analytic = result["gradients"][:, 0]
    scale = np.abs(analytic).max()
    checked = 0
    for field in range(58):
        if field >= 10 and (field - 10) % 16 >= (degree + 1) ** 2:
            assert analytic[field] == 0
            continue
        h = 1e-3 if field < 10 else 1e-3
        high, low = params.copy(), params.copy()
        high[field, 0] += np.float32(h)
        low[field, 0] -= np.float32(h)
        step = float(high[field, 0]) - float(low[field, 0])
        fd = (oracle.loss_f64(oracle.render_reference(high, cam, (0.1, 0.1, 0.1), degree), target)
              - oracle.loss_f64(oracle.render_reference(low, cam, (0.1, 0.1, 0.1), degree), target)) / step
        assert abs(fd - analytic[field]) <= 2e-4 * abs(analytic[field]) + 1e-6 * scale, f"field {field}"
        checked += 1
    assert checked == 10 + 3 * (degree + 1) ** 2


def test_statistic_definition(oracle):
    """statistics[0] = |(dmean_x W/2, dmean_y H/2)|, observed flag, absolute row."""
    cam = _camera(oracle, 40, 30)
    screen = np.zeros((11, 2))
    screen[0, 0], screen[1, 0], screen[9, 0], screen[10, 0] = 0.3, -0.4, 0.5, 0.7
    params = _scenes().make_gaussians(2)
    mapped = oracle.map_f64(params, cam, 0, screen, np.array([1, 0], np.uint8))
    assert abs(mapped["statistics"][0, 0] - math.hypot(0.3 * 20, -0.4 * 15)) < 1e-12
    assert mapped["statistics"][1].tolist() == [1.0, 0.0]
    assert abs(mapped["absolute_statistics"][0, 0] - math.hypot(0.5 * 20, 0.7 * 15)) < 1e-12


def test_tile_lists_match_brute_force_ordering(oracle):
    scenes = _scenes()
    params = scenes.config0_params(seed=3, count=3000)
    cam = _camera(oracle, 256, 256, None, (0.0, 0.0, 3.0))
    cam.fx = cam.fy = 221.7025
    projected = oracle.project_f32(params, cam, 0)
    lists = oracle.tile_lists(cam, projected)
    keys, ids = [], []
    for g in range(3000):
        if projected["record_counts"][g] == 0:
            continue
        x0, y0, x1, y1 = projected["tile_rects"][g]
        depth = int(projected["depths"][g:g + 1].view(np.uint32)[0])
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                keys.append(((ty * cam.tile_count_x + tx) << 32) | depth)
                ids.append(g)
    keys, ids = np.array(keys, np.uint64), np.array(ids, np.int32)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(lists["sorted_keys"], keys[order])
    np.testing.assert_array_equal(lists["tile_gaussian_ids"], ids[order])
    tiles = (keys[order] >> np.uint64(32)).astype(np.int64)
    np.testing.assert_array_equal(lists["tile_offsets"], np.searchsorted(tiles, np.arange(cam.tile_count_x * cam.tile_count_y + 1)))


def test_project_f32_close_to_f64(oracle):
    scenes = _scenes()
    params = scenes.config0_params(seed=4, count=2000, sh_degree=3)
    cam = _camera(oracle, 256, 256, None, (0.0, 0.0, 3.0))
    cam.fx = cam.fy = 221.7025
    a = oracle.project_f32(params, cam, 3)
    b = oracle.project_f64(params, cam, 3)
    visible = (a["record_counts"][:-1] > 0) & (b["visible"] > 0)
    assert visible.sum() > 1000
    np.testing.assert_allclose(a["means"][visible], b["means"][visible], rtol=1e-5)
    np.testing.assert_allclose(a["conic_opacities"][visible], b["conics"][visible], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(a["colors"][visible], b["colors"][visible], rtol=1e-5, atol=1e-6)
    assert np.mean(np.all(a["tile_rects"][visible] == b["tile_rects"][visible], axis=1)) > 0.999


def test_render_f32_close_to_f64(oracle):
    scenes = _scenes()
    params = scenes.config0_params(seed=5, count=3000)
    cam = _camera(oracle, 128, 128, None, (0.0, 0.0, 3.0))
    cam.fx = cam.fy = 110.85
    projected = oracle.project_f32(params, cam, 0)
    lists = oracle.tile_lists(cam, projected)
    image32, ambiguous = oracle.render_f32(cam, projected, lists, (0.0, 0.0, 0.0))
    reference = oracle.project_f64(params, cam, 0)
    reference["tile_rects"] = projected["tile_rects"]
    reference["visible"] = ((projected["record_counts"][:-1] > 0) & (reference["visible"] > 0)).astype(np.uint8)
    image64, ambiguous64 = oracle.render_f64(cam, reference, (0.0, 0.0, 0.0))
    clear = (ambiguous == 0) & (ambiguous64 == 0)
    assert clear.mean() > 0.98
    assert np.abs(image32 - image64)[clear].max() < 1e-4


def test_adam_first_step_closed_form(oracle):
    """solver_test.cu:218-219: with zeroed moments at step 2 the update is lr * 0.1 c1 / sqrt(0.001 c2)."""
    g = np.full((58, 1), 0.37, np.float32)
    params = np.zeros((58, 1), np.float32)
    zero = np.zeros((58, 1), np.float32)
    c1 = np.float32(1.0 / (1.0 - 0.9 ** 2))
    c2 = np.float32(1.0 / (1.0 - 0.999 ** 2))
    out, m, v = oracle.adam_f32(g, params, zero, zero, 1e-4, c1, c2)
    factor = (0.1 / (1 - 0.9 ** 2)) / math.sqrt(0.001 / (1 - 0.999 ** 2))
    assert abs(-out[9, 0] - 0.025 * factor) < GOLDEN["adam_reset_step"]["tolerance"]
    assert abs(-out[0, 0] - 1e-4 * factor) < 1e-9
    assert abs(-out[11, 0] - 0.000125 * factor) < 1e-9


def test_philox_known_answers(oracle):
    """Philox4x32-10 known-answer vectors published with Random123 (Salmon et al., SC'11)."""
    import ctypes
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for counter, key, expected in cases:
        block = (ctypes.c_uint32 * 4)(*counter)
        oracle.lib.oracle_philox4x32_10(block, key[0], key[1])
        assert tuple(block) == expected


def test_jitter_normals_are_standard(oracle):
    z = oracle.jitter_normals(42, 3, 40000).astype(np.float64)
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.03
    assert np.all(np.isfinite(z))


def _apply(oracle, params, sums, abs_sums, counts, threshold=0.15, max_screen=0.0, absolute=0.0, seed=42, counter=0):
    decisions, offsets = oracle.density_count_f32(params, sums, abs_sums, counts, threshold, np.float32(0.01), 0.005,
                                                  np.float32(0.1), max_screen > 0, absolute)
    out, history, parents, flags = oracle.density_write_f32(params, decisions, offsets, seed, counter)
    n = params.shape[1]
    totals = offsets[n]
    return out, history, parents, (int(totals[1] + 2 * totals[3]), int(n - totals[0]))


def test_density_reference_vectors(oracle):
    """density_test.cu literal vectors through the oracle (statistics accumulated as in the
    reference's AccumulateRoundStatisticsKernel)."""
    scenes = _scenes()
    d = GOLDEN["density"]
    params = scenes.make_gaussians(3)
    params[9, 1] = np.log(np.float32(0.001) / np.float32(0.999))
    z3 = np.zeros(3, np.float32)
    out, history, _, changes = _apply(oracle, params, z3, z3, np.zeros(3, np.uint32))
    assert history.tolist() == d["low_opacity"]["parents"] and list(changes) == d["low_opacity"]["changes"]
    params = scenes.make_gaussians(6)
    for i in range(6):
        params[3:6, i] = np.log(np.float32(0.005 if i < 2 else 0.02))
    params[3:6, 5] = np.log(np.float32(0.01))
    sums = np.array([0.0003, 0.0001, 0.0003, 0.0001, 0.0003, 0.0003], np.float32)
    abs_sums = np.array([0.001, 0.001, 0.0007, 0.001, 0.001, 0.0007], np.float32)
    out, history, parents, _ = _apply(oracle, params, sums, abs_sums, np.ones(6, np.uint32), threshold=0.0002,
                                      absolute=0.0008)
    assert parents.tolist() == d["signed_vs_absolute"]["parent_ids"]
    assert history.tolist() == d["signed_vs_absolute"]["history"]
    params = scenes.make_gaussians(1)
    params[3:6] = np.log(np.float32(0.2))
    out, history, parents, _ = _apply(oracle, params, np.ones(1, np.float32), np.ones(1, np.float32),
                                      np.ones(1, np.uint32))
    assert history.tolist() == d["split"]["history"] and parents.tolist() == d["split"]["parent_ids"]
    np.testing.assert_allclose(np.exp(out[3:6]), d["split"]["child_scale"], atol=1e-7)
    for threshold, count in zip(d["round_mean"]["thresholds"], d["round_mean"]["counts"]):
        params = scenes.make_gaussians(1)
        params[3:6] = np.log(np.float32(0.005))
        sums = np.array([np.float32(0.1) + np.float32(0.3)], np.float32)
        out, _, _, _ = _apply(oracle, params, sums, sums, np.array([2], np.uint32), threshold=threshold)
        assert out.shape[1] == count


def test_knn_oracle_matches_float64_brute_force(oracle):
    rng = np.random.default_rng(17)
    for count in GOLDEN["knn_sizes"]:
        points = rng.uniform(-10, 10, size=(count, 3)).astype(np.float32)
        neighbors, distances = oracle.knn3(points)
        for i in range(count):
            d = ((points[i].astype(np.float64) - points.astype(np.float64)) ** 2).sum(1)
            d[i] = np.inf
            order = np.lexsort((np.arange(count), d))[: min(3, count - 1)]
            value = d[order].mean() if len(order) else 0.0
            assert abs(distances[i] - value) <= GOLDEN["knn_relative"] * max(1.0, value)
            assert neighbors[i][: len(order)].tolist() == order.tolist()
