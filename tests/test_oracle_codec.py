"""Oracle s4-bp128-d4 codec pinned byte-for-byte by the reference codec's
output (tests/golden/codec_s4bp128d4.npz, written by cpu_2d/compression/cpusimd
compiled from source), plus the CpuSimd gate (cpusimd.hh:71-173)."""
import numpy as np
import pytest

import oracle_lib as O

GOLD = np.load(O.golden("codec_s4bp128d4.npz"))
CASES = sorted(k[3:] for k in GOLD.files if k.startswith("in_"))


@pytest.mark.parametrize("case", CASES)
def test_encode_matches_reference_stream(case):
    x = GOLD["in_" + case]
    want = GOLD["out_" + case]
    got = O.s4_encode(x)
    assert got.dtype == np.uint32
    assert np.array_equal(got, want)


@pytest.mark.parametrize("case", CASES)
def test_decode_reference_stream(case):
    x = GOLD["in_" + case]
    rc, got = O.s4_decode(GOLD["out_" + case], len(x))
    assert rc == 0 and np.array_equal(got, x)


def test_header_layout():
    w = O.s4_encode(np.arange(300, dtype=np.uint32) * 3)
    assert w[0] == 256 and list(w[1:4]) == [123456] * 3
    # one partial group: widths of 2 miniblocks then zeros (deltas are 12 => 4 bits)
    assert w[4] == (4 << 24) | (4 << 16) and w[5] == 0 and w[6] == 0 and w[7] == 0


def test_max_words_bound_holds():
    rng = np.random.default_rng(3)
    for n in (1, 127, 128, 129, 5000):
        x = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        assert len(O.s4_encode(x)) <= O.lib().orc_s4bp128d4_max_words(O.c_sz(n))


def test_cpusimd_gate():
    q = np.arange(0, 640, 5, dtype=np.int64)  # 128 values
    raw = O.cpusimd_compress(q, threshold=200)
    assert np.array_equal(raw, q.astype(np.uint32))
    assert np.array_equal(O.cpusimd_decompress(raw, len(q), 200), q)
    comp = O.cpusimd_compress(q, threshold=64)
    assert np.array_equal(comp, O.s4_encode(q.astype(np.uint32)))
    assert len(comp) != len(q)
    assert np.array_equal(O.cpusimd_decompress(comp, len(q), 64), q)


def test_cpusimd_h2_rule_reproduced():
    # A compressed stream whose length equals n is treated as raw by the
    # receiver (cpusimd.hh:170-173).  Two miniblocks with constant 4-strided
    # delta 2^30 (31 bits) give 8 + 2*124 = 256 words; four tail values with
    # 4-byte varints add 4 words: 260 words for n = 260.
    k = np.arange(256)
    x = ((k // 4) % 4) * (1 << 30)
    tail = np.array([1 << 22, 2 << 22, 3 << 22, 4 << 22])
    q = np.concatenate([x, tail]).astype(np.int64)
    c = O.cpusimd_compress(q, 64)
    assert len(c) == len(q) == 260
    out = O.cpusimd_decompress(c, len(q), 64)
    assert np.array_equal(out, c.astype(np.int64))  # mis-decoded as raw, as in the reference
