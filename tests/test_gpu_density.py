"""Density control and KNN parity on the GPU: bit-exact masks, compaction and split outputs
against the oracle, the reference's own density and KNN cases (test/density_test.cu,
test/knn_test.cu) with their literal vectors, and the split-jitter generator."""
import ctypes
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
DENSITY = GOLDEN["density"]


class Density:
    """DensityController semantics (density_controller.cu:410-647) over the C ABI."""

    def __init__(self, gsn, context, count):
        self.gsn, self.context = gsn, context
        n = max(count, 1)
        self.sums = gsn.DeviceBuffer.zeros((n,), np.float32)
        self.abs_sums = gsn.DeviceBuffer.zeros((n,), np.float32)
        self.counts = gsn.DeviceBuffer.zeros((n,), np.uint32)
        self.capacity = n
        self.seed = 0
        self.counter = 0

    def accumulate(self, statistics, absolute, count):
        s = self.gsn.DeviceBuffer.from_numpy(np.asarray(statistics, np.float32))
        a = self.gsn.DeviceBuffer.from_numpy(np.asarray(absolute, np.float32))
        self.gsn.call("gsn_density_accumulate", self.context.handle, s.ptr, a.ptr, count, self.sums.ptr,
                      self.abs_sums.ptr, self.counts.ptr)

    def accumulate_observations(self, gradients, radii, ids, count):
        g = self.gsn.DeviceBuffer.from_numpy(np.asarray(gradients, np.float32))
        r = self.gsn.DeviceBuffer.from_numpy(np.asarray(radii, np.float32))
        i = self.gsn.DeviceBuffer.from_numpy(np.asarray(ids, np.int32))
        self.gsn.call("gsn_density_accumulate_observations", self.context.handle, g.ptr, r.ptr, i.ptr, len(radii),
                      count, self.sums.ptr, self.counts.ptr)

    def reserve(self, count):
        """Grows the accumulators to `count` entries, keeping their values (Reserve)."""
        if count <= self.capacity:
            return
        gsn = self.gsn
        keep = [buffer.numpy() for buffer in (self.sums, self.abs_sums, self.counts)]
        grown = []
        for old in keep:
            values = np.zeros(count, old.dtype)
            values[: old.shape[0]] = old
            grown.append(gsn.DeviceBuffer.from_numpy(values))
        self.sums, self.abs_sums, self.counts = grown
        self.capacity = count

    def control(self, params, extent=1.0, threshold=0.15, percent=0.01, min_opacity=0.005, max_screen=0.0, seed=42,
                absolute_threshold=0.0):
        gsn = self.gsn
        count = params.shape[1]
        self.reserve(count)
        if seed != self.seed:
            self.seed, self.counter = seed, 0
        p = gsn.DeviceBuffer.from_numpy(params) if count else gsn.DeviceBuffer((1,), np.float32)
        decisions = gsn.DeviceBuffer((count + 1, 4), np.int32)
        offsets = gsn.DeviceBuffer((count + 1, 4), np.int32)
        totals = (ctypes.c_int32 * 4)()
        gsn.call("gsn_density_count", self.context.handle, p.ptr, count, self.sums.ptr, self.abs_sums.ptr,
                 self.counts.ptr, threshold, np.float32(percent) * np.float32(extent), min_opacity,
                 np.float32(0.1) * np.float32(extent), int(max_screen > 0), absolute_threshold, decisions.ptr,
                 offsets.ptr, totals)
        t = list(totals)
        out_count = t[0] + t[1] + 2 * t[3]
        out = gsn.DeviceBuffer((58, max(out_count, 1)), np.float32)
        history = gsn.DeviceBuffer((max(out_count, 1),), np.int32)
        parents = gsn.DeviceBuffer((max(out_count, 1),), np.int32)
        flags = gsn.DeviceBuffer((max(out_count, 1),), np.float32)
        gsn.call("gsn_density_write", self.context.handle, p.ptr, count, decisions.ptr, offsets.ptr, totals, self.seed,
                 self.counter, out.ptr, out_count, history.ptr, parents.ptr, flags.ptr)
        gsn.call("gsn_density_clear", self.context.handle, self.sums.ptr, self.abs_sums.ptr, self.counts.ptr, count)
        self.counter += 1
        self.context.synchronize()
        self.reserve(out_count)
        return {"params": out.numpy()[:, :out_count], "count": out_count, "history": history.numpy()[:out_count],
                "parents": parents.numpy()[:out_count], "flags": flags.numpy()[:out_count],
                "changes": (t[1] + 2 * t[3], count - t[0]), "decisions": decisions.numpy(), "offsets": offsets.numpy()}


def test_density_bit_exact_against_oracle(gsn, context, oracle, scenes):
    rng = np.random.default_rng(12)
    n = 50_000
    params = scenes.config0_params(seed=13, count=n, sh_degree=3)
    params[3:6] = rng.normal(-4.5, 1.0, size=(3, n)).astype(np.float32)
    sums = rng.exponential(2e-4, size=n).astype(np.float32) * 5
    abs_sums = (sums * rng.uniform(1.0, 3.0, size=n)).astype(np.float32)
    counts = rng.integers(0, 6, size=n).astype(np.uint32)
    for prune_large, absolute in ((False, 0.0), (True, 8e-4)):
        density = Density(gsn, context, n)
        density.sums.upload(sums)
        density.abs_sums.upload(abs_sums)
        density.counts.upload(counts)
        result = density.control(params, extent=1.7, threshold=2e-4, max_screen=20.0 if prune_large else 0.0,
                                 seed=5, absolute_threshold=absolute)
        decisions, offsets = oracle.density_count_f32(params, sums, abs_sums, counts, 2e-4,
                                                      np.float32(0.01) * np.float32(1.7), 0.005,
                                                      np.float32(0.1) * np.float32(1.7), prune_large, absolute)
        np.testing.assert_array_equal(result["decisions"], decisions)
        np.testing.assert_array_equal(result["offsets"], offsets)
        out, history, parents, flags = oracle.density_write_f32(params, decisions, offsets, 5, density.counter - 1)
        assert result["count"] == out.shape[1]
        assert decisions[:, 3].sum() > 10 and decisions[:, 1].sum() > 10
        np.testing.assert_array_equal(result["params"].view(np.uint32), out.view(np.uint32))
        np.testing.assert_array_equal(result["history"], history)
        np.testing.assert_array_equal(result["parents"], parents)
        np.testing.assert_array_equal(result["flags"], flags)


def test_jitter_generator_matches_oracle(gsn, context, oracle):
    out = gsn.DeviceBuffer((4099,), np.float32)
    gsn.call("gsn_density_normals", context.handle, 1234, 7, 4099, out.ptr)
    context.synchronize()
    np.testing.assert_array_equal(out.numpy().view(np.uint32), oracle.jitter_normals(1234, 7, 4099).view(np.uint32))


def test_reference_case_low_opacity_removed(gsn, context, scenes):
    """density_test.cu:55-73."""
    params = scenes.make_gaussians(3)
    params[9, 1] = np.log(np.float32(0.001) / np.float32(0.999))
    result = Density(gsn, context, 3).control(params)
    case = DENSITY["low_opacity"]
    assert result["count"] == 2
    assert list(result["history"]) == case["parents"]
    assert list(result["changes"]) == case["changes"]
    np.testing.assert_array_equal(result["params"][:, 0], params[:, 0])
    np.testing.assert_array_equal(result["params"][:, 1], params[:, 2])


def test_reference_case_clone_and_clear(gsn, context, scenes):
    """density_test.cu:75-95."""
    params = scenes.make_gaussians(3)
    params[3:6] = np.log(np.float32(0.005))
    density = Density(gsn, context, 3)
    density.accumulate_observations([0.12, 0.16, 50, 50, 0, 0], [1, 0, 1], [0, 0, 1], 3)
    first = density.control(params)
    case = DENSITY["clone_then_clear"]
    assert first["count"] == case["count"]
    assert list(first["history"]) == case["parents"]
    np.testing.assert_array_equal(first["params"][:, 0], first["params"][:, 3])
    second = density.control(first["params"])
    assert second["count"] == case["count"]
    assert second["changes"][0] == case["second_changes"]


def test_reference_case_split_reproducible(gsn, context, scenes):
    """density_test.cu:97-130."""
    params = scenes.make_gaussians(1)
    params[3:6] = np.log(np.float32(0.2))
    results = []
    for _ in range(2):
        density = Density(gsn, context, 1)
        density.accumulate([1, 1], [1, 1], 1)
        results.append(density.control(params))
    first, second = results
    case = DENSITY["split"]
    assert first["count"] == 2
    assert list(first["history"]) == case["history"]
    assert list(first["parents"]) == case["parent_ids"]
    np.testing.assert_array_equal(first["params"], second["params"])
    actual = first["params"]
    assert np.all(np.isfinite(actual))
    np.testing.assert_allclose(np.exp(actual[3:6]), case["child_scale"], atol=1e-7)
    np.testing.assert_array_equal(actual[6:], np.repeat(params[6:], 2, axis=1))
    assert np.abs(actual[0:3, 0] - actual[0:3, 1]).sum() > 1e-6


@pytest.mark.parametrize("threshold", DENSITY["round_mean"]["thresholds"])
def test_reference_case_round_mean(gsn, context, scenes, threshold):
    """density_test.cu:132-150: the mean over accumulated rounds is compared with the threshold."""
    params = scenes.make_gaussians(1)
    params[3:6] = np.log(np.float32(0.005))
    density = Density(gsn, context, 1)
    density.accumulate([0.1, 1], [0.1, 1], 1)
    density.accumulate([0.3, 1], [0.3, 1], 1)
    result = density.control(params, threshold=threshold)
    case = DENSITY["round_mean"]
    index = case["thresholds"].index(threshold)
    assert result["count"] == case["counts"][index]
    assert result["changes"][0] == case["changes"][index]
    assert result["changes"][1] == 0


def test_reference_case_oversize_prune(gsn, context, scenes):
    """density_test.cu:152-163."""
    params = scenes.make_gaussians(2)
    params[3:6, 0] = np.log(np.float32(0.2))
    params[3:6, 1] = np.log(np.float32(0.01))
    result = Density(gsn, context, 2).control(params, threshold=0.15, max_screen=20)
    assert result["count"] == 1
    assert list(result["history"]) == DENSITY["oversize"]["parents"]


@pytest.mark.parametrize("cap", DENSITY["opacity_caps"])
def test_reference_case_opacity_reset(gsn, context, scenes, cap):
    """density_test.cu:165-194."""
    params = scenes.make_gaussians(4)
    params[9] = [np.log(np.float32(0.95) / np.float32(0.05)), np.log(np.float32(0.2) / np.float32(0.8)),
                 np.log(np.float32(0.8) / np.float32(0.2)), -9.0]
    device = gsn.DeviceBuffer.from_numpy(params)
    limit = np.float32(np.log(np.float32(cap) / (np.float32(1.0) - np.float32(cap))))
    gsn.call("gsn_density_reset_opacity", context.handle, device.ptr, 4, 4, float(limit))
    actual = device.numpy()
    expected = params.copy()
    expected[9] = np.minimum(params[9], limit)
    np.testing.assert_array_equal(actual, expected)


def test_reference_case_all_pruned_and_empty(gsn, context, scenes):
    """density_test.cu:196-209."""
    params = scenes.make_gaussians(2)
    params[9] = -12.0
    density = Density(gsn, context, 2)
    assert density.control(params)["count"] == DENSITY["all_pruned"]["count"]
    assert density.control(params[:, :0])["count"] == 0


def test_reference_case_signed_and_absolute(gsn, context, scenes):
    """density_test.cu:210-235: signed gradients select clones, absolute ones select splits."""
    params = scenes.make_gaussians(6)
    for i in range(6):
        params[3:6, i] = np.log(np.float32(0.005 if i < 2 else 0.02))
    params[3:6, 5] = np.log(np.float32(0.01))
    density = Density(gsn, context, 6)
    density.accumulate([0.0003, 0.0001, 0.0003, 0.0001, 0.0003, 0.0003, 1, 1, 1, 1, 1, 1],
                       [0.001, 0.001, 0.0007, 0.001, 0.001, 0.0007], 6)
    result = density.control(params, threshold=0.0002, absolute_threshold=0.0008)
    case = DENSITY["signed_vs_absolute"]
    assert result["count"] == case["count"]
    assert list(result["parents"]) == case["parent_ids"]
    assert list(result["history"]) == case["history"]
    following = density.control(result["params"], threshold=0.0002, absolute_threshold=0.0008)
    assert following["count"] == case["count"]
    assert following["changes"][0] == 0


@pytest.mark.parametrize("threshold", DENSITY["absolute_mean"]["thresholds"])
def test_reference_case_absolute_mean(gsn, context, scenes, threshold):
    """density_test.cu:237-254."""
    params = scenes.make_gaussians(1)
    params[3:6] = np.log(np.float32(0.02))
    density = Density(gsn, context, 1)
    density.accumulate([0.0, 1.0], [0.0007], 1)
    density.accumulate([0.0, 1.0], [0.001], 1)
    result = density.control(params, threshold=0.0002, absolute_threshold=threshold)
    case = DENSITY["absolute_mean"]
    assert result["count"] == case["counts"][case["thresholds"].index(threshold)]


# ---- KNN ------------------------------------------------------------------------------------

@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_reference_case_knn_exhaustive(gsn, context, oracle, mode):
    """knn_test.cu:8-42 (degenerate clouds), plus bit-exact neighbours against the oracle."""
    rng = np.random.default_rng(17 + mode)
    for count in GOLDEN["knn_sizes"]:
        points = rng.uniform(-10, 10, size=(count, 3)).astype(np.float32)
        if mode == 1:
            points[:, 0] = 0
        if mode == 2:
            points[:] = 0
        if mode == 3:
            points[:, 1:] = 0
        neighbors, distances = gsn.knn3(context, points)
        expected_neighbors, expected = oracle.knn3(points)
        np.testing.assert_array_equal(neighbors, expected_neighbors)
        np.testing.assert_array_equal(distances.view(np.uint32), expected.view(np.uint32))
        for i in range(count):
            d = np.sort(((points[i].astype(np.float64) - points.astype(np.float64)) ** 2).sum(1))[1:]
            k = min(3, len(d))
            value = d[:k].mean() if k else 0.0
            if value == 0:
                assert distances[i] == 0
            else:
                assert abs(distances[i] - value) <= GOLDEN["knn_relative"] * max(1.0, value)


def test_reference_case_knn_order(gsn, context):
    """knn_test.cu:44-53."""
    points = np.array(GOLDEN["knn_order_points"], np.float32)
    _, first = gsn.knn3(context, points)
    _, second = gsn.knn3(context, points[::-1].copy())
    np.testing.assert_array_equal(first, second[::-1])


def test_knn_large_bit_exact(gsn, context, oracle, scenes):
    params = scenes.blob_scene(20_000, seed=2)
    points = params[0:3].T.copy()
    points[:100] = points[100:200]  # duplicates exercise the index tie rule
    neighbors, distances = gsn.knn3(context, points)
    expected_neighbors, expected = oracle.knn3(points)
    np.testing.assert_array_equal(neighbors, expected_neighbors)
    np.testing.assert_array_equal(distances.view(np.uint32), expected.view(np.uint32))
