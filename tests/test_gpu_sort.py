"""The stable LSD radix sort (per pass: per-tile digit histograms, a device scan of the histogram
table, a ranking scatter over 4096-element tiles) against NumPy's stable argsort, through
gsn_selftest_sort_pairs: sizes around the tile boundary, many tiles, every digit width the
library uses (13-bit tile keys, 30-bit Morton codes, 32-bit depth keys), keys that all share one
digit, and keys whose digits all differ; equal keys keep their input order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sort(gsn, context, keys, values, bits):
    n = keys.shape[0]
    k = gsn.DeviceBuffer.from_numpy(keys.astype(np.uint32))
    v = gsn.DeviceBuffer.from_numpy(values.astype(np.uint32))
    ko = gsn.DeviceBuffer((max(n, 1),), np.uint32)
    vo = gsn.DeviceBuffer((max(n, 1),), np.uint32)
    gsn.call("gsn_selftest_sort_pairs", context.handle, k.ptr, v.ptr, n, bits, ko.ptr, vo.ptr)
    context.synchronize()
    return ko.numpy()[:n], vo.numpy()[:n]


def _expected(keys, values, bits):
    masked = keys.astype(np.uint64) & ((1 << bits) - 1)
    order = np.argsort(masked, kind="stable")
    return keys[order], values[order]


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 12289, 1_000_003])
@pytest.mark.parametrize("bits", [13, 30, 32])
def test_sort_matches_stable_argsort(gsn, context, n, bits):
    rng = np.random.default_rng(n * 100 + bits)
    keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
    keys[rng.random(n) < 0.3] = keys[0]  # many equal keys: stability
    values = np.arange(n, dtype=np.uint32)
    got_keys, got_values = _sort(gsn, context, keys, values, bits)
    want_keys, want_values = _expected(keys, values, bits)
    np.testing.assert_array_equal(got_keys, want_keys)
    np.testing.assert_array_equal(got_values, want_values)


def test_sort_uses_only_the_low_bits(gsn, context):
    rng = np.random.default_rng(5)
    n = 50_000
    keys = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    values = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    got_keys, got_values = _sort(gsn, context, keys, values, 13)
    want_keys, want_values = _expected(keys, values, 13)
    np.testing.assert_array_equal(got_keys, want_keys)
    np.testing.assert_array_equal(got_values, want_values)


@pytest.mark.parametrize("pattern", ["equal", "descending", "one_digit_each"])
def test_sort_degenerate_keys(gsn, context, pattern):
    n = 3 * 4096 + 17
    if pattern == "equal":
        keys = np.full(n, 0xdeadbeef, np.uint32)
    elif pattern == "descending":
        keys = np.arange(n, 0, -1, dtype=np.uint32)
    else:
        keys = (np.arange(n, dtype=np.uint32) % 256) * 0x01010101
    values = np.arange(n, dtype=np.uint32)
    got_keys, got_values = _sort(gsn, context, keys, values, 32)
    want_keys, want_values = _expected(keys, values, 32)
    np.testing.assert_array_equal(got_keys, want_keys)
    np.testing.assert_array_equal(got_values, want_values)


@pytest.mark.timeout(300)
def test_sort_at_tile_list_size(gsn, context):
    """30M 13-bit keys, two passes over 7,325 tiles (the record count of a 3M-Gaussian 1080p view)."""
    rng = np.random.default_rng(13)
    n = 30_000_000
    keys = np.sort(rng.integers(0, 8160, size=n // 8, dtype=np.uint32))  # locally clustered, as emitted
    keys = np.repeat(keys, 8)[:n] ^ rng.integers(0, 2, size=n, dtype=np.uint32)
    values = np.arange(n, dtype=np.uint32)
    got_keys, got_values = _sort(gsn, context, keys, values, 13)
    want_keys, want_values = _expected(keys, values, 13)
    np.testing.assert_array_equal(got_keys, want_keys)
    np.testing.assert_array_equal(got_values, want_values)
