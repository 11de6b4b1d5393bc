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
