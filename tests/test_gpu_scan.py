"""Device exclusive scan (used for record offsets, depth-order record counts and the radix-sort
histograms) against numpy's cumulative sum, through gsn_scan_record_offsets, which scans
count + 1 ints in place and returns the total: single-block inputs, block boundaries, the
largest input whose block totals each block sums itself (2048 blocks of 4096), the first one
that scans the totals recursively, and a pointer that is not 16-byte aligned (scalar loads)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan(gsn, context, values: np.ndarray, offset_ints: int = 0) -> tuple[np.ndarray, int]:
    padded = np.concatenate([np.full(offset_ints, 7, np.int32), values.astype(np.int32)])
    buffer = gsn.DeviceBuffer.from_numpy(padded)
    total = ctypes.c_int()
    count = values.size - 1
    pointer = ctypes.c_void_p(buffer.ptr.value + 4 * offset_ints)
    gsn.call("gsn_scan_record_offsets", context.handle, pointer, count, ctypes.byref(total))
    context.synchronize()
    result = buffer.numpy()
    assert np.all(result[:offset_ints] == 7), "scan wrote before its input"
    return result[offset_ints:], total.value


@pytest.mark.parametrize("entries", [1, 2, 4095, 4096, 4097, 3_000_001, 2048 * 4096, 2048 * 4096 + 1])
def test_exclusive_scan_matches_cumsum(gsn, context, entries):
    rng = np.random.default_rng(entries)
    values = rng.integers(0, 100, entries).astype(np.int32)
    print(f"scan of {entries} entries", flush=True)
    result, total = _scan(gsn, context, values)
    expected = np.concatenate([[0], np.cumsum(values[:-1], dtype=np.int64)])
    np.testing.assert_array_equal(result, expected)
    assert total == expected[-1]


@pytest.mark.parametrize("entries", [4097, 1_000_003])
def test_exclusive_scan_unaligned_pointer(gsn, context, entries):
    rng = np.random.default_rng(entries + 1)
    values = rng.integers(0, 1000, entries).astype(np.int32)
    result, total = _scan(gsn, context, values, offset_ints=1)
    expected = np.concatenate([[0], np.cumsum(values[:-1], dtype=np.int64)])
    np.testing.assert_array_equal(result, expected)
    assert total == expected[-1]
