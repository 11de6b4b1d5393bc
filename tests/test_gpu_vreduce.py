"""vreduce combines on the device vs the oracle's closed forms."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,n", [(1, 10), (4, 262144), (4, 1001), (3, 7)])
def test_or(k, n):
    rng = np.random.default_rng(k * n)
    maps = rng.integers(0, 1 << 63, size=(k, n), dtype=np.uint64) & rng.integers(0, 1 << 63, size=(k, n), dtype=np.uint64)
    assert np.array_equal(g5.vreduce_or(maps), O.vreduce_or(maps))


@pytest.mark.parametrize("k,n", [(1, 10), (4, 1 << 20), (4, 999), (2, 5)])
def test_min_pred(k, n):
    rng = np.random.default_rng(k + n)
    preds = rng.integers(-1, 1 << 24, size=(k, n)).astype(np.int64)
    preds[rng.random((k, n)) < 0.5] = -1
    assert np.array_equal(g5.vreduce_min_pred(preds), O.vreduce_min_pred(preds))


@pytest.mark.parametrize("k,n", [(1, 10), (4, 262144), (4, 1001), (16, 4096), (64, 130)])
def test_list_forms(k, n):
    """The pointer-list forms equal the contiguous ones (buffers anywhere)."""
    rng = np.random.default_rng(3 * k + n)
    maps = rng.integers(0, 1 << 63, size=(k, n), dtype=np.uint64) & rng.integers(0, 1 << 63, size=(k, n), dtype=np.uint64)
    assert np.array_equal(g5.vreduce_or_list(list(maps)), O.vreduce_or(maps))
    preds = rng.integers(-1, 1 << 24, size=(k, n)).astype(np.int64)
    preds[rng.random((k, n)) < 0.5] = -1
    assert np.array_equal(g5.vreduce_min_pred_list(list(preds)), O.vreduce_min_pred(preds))


def test_list_forms_reject_bad_arguments():
    a = g5.DeviceArray(8, np.uint64)
    o = g5.DeviceArray(8, np.uint64)
    ptrs = (g5.ctypes.c_void_p * 65)(*([a.ptr] * 65))
    with pytest.raises(g5.G5Error):
        g5.check(g5.lib().g5_vreduce_or_list(ptrs, 65, 8, o.ptr))   # more than 64 maps
    with pytest.raises(g5.G5Error):
        g5.check(g5.lib().g5_vreduce_or_list(ptrs, 0, 8, o.ptr))
