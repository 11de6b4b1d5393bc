"""Device s4-bp128-d4 codec: word streams identical to the reference codec's
(golden fixtures written by the compiled reference), lossless round trips,
and malformed streams rejected."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu

GOLD = np.load(O.golden("codec_s4bp128d4.npz"))
CASES = sorted(k[3:] for k in GOLD.files if k.startswith("in_"))


@pytest.mark.parametrize("case", CASES)
def test_encode_matches_reference(case):
    assert np.array_equal(g5.s4_encode(GOLD["in_" + case]), GOLD["out_" + case])


@pytest.mark.parametrize("case", CASES)
def test_decode_reference_stream(case):
    x = GOLD["in_" + case]
    assert np.array_equal(g5.s4_decode(GOLD["out_" + case], len(x)), x)


@pytest.mark.parametrize("n", [0, 3, 128, 1000, 2048, 2049, 40000, 300001])
def test_random_round_trip_and_oracle_bytes(n):
    rng = np.random.default_rng(n)
    for kind in ("sorted", "random"):
        if kind == "sorted":
            x = np.sort(rng.choice(1 << 26, size=n, replace=False)).astype(np.uint32) if n else np.zeros(0, np.uint32)
        else:
            x = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        w = g5.s4_encode(x)
        assert np.array_equal(w, O.s4_encode(x))
        assert np.array_equal(g5.s4_decode(w, n), x)


def test_large_sorted_queue():
    rng = np.random.default_rng(99)
    n = 1 << 22
    x = np.sort(rng.choice(1 << 24, size=n, replace=False)).astype(np.uint32)
    w = g5.s4_encode(x)
    assert np.array_equal(w, O.s4_encode(x))
    assert np.array_equal(g5.s4_decode(w, n), x)


def test_malformed_streams_rejected():
    x = np.arange(0, 3000, 3, dtype=np.uint32)
    w = g5.s4_encode(x)
    bad = w.copy()
    bad[1] = 7  # cookie
    with pytest.raises(g5.G5Error):
        g5.s4_decode(bad, len(x))
    with pytest.raises(g5.G5Error):
        g5.s4_decode(w[:6], len(x))  # truncated
    with pytest.raises(g5.G5Error):
        g5.s4_decode(w, 10)  # output capacity too small
