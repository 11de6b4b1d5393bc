"""g5_s4_roundtrip_noalloc, the codec round trip that allocates nothing (its
scratch is a static device array), against g5_s4_roundtrip, the allocating
entry point it mirrors: the same stream words, stream length and decoded
values on consecutive calls, the second smaller than the first (the array
keeps the last call's contents), at the size limit, and the limit's error.
The streams are also the oracle encoder's (the reference codec's format)."""
import ctypes

import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu

MAX = 1 << 22


def _round_trip(fn, x):
    din = g5.DeviceArray.from_host(x)
    cap = g5.s4_max_words(x.size)
    dw = g5.DeviceArray(cap, np.uint32)
    dout = g5.DeviceArray(max(x.size, 1), np.uint32)
    nw = g5.c_sz()
    g5.check(fn(din.ptr, x.size, dw.ptr, cap, dout.ptr, ctypes.byref(nw)))
    return nw.value, dw.to_host(nw.value), dout.to_host(x.size)


def _queue(rng, n, hi):
    return np.sort(rng.choice(hi, size=n, replace=False)).astype(np.uint32)


def test_noalloc_equals_allocating_on_consecutive_calls():
    g5.init(0)
    L = g5.lib()
    rng = np.random.default_rng(11)
    # multi-kernel path, then a smaller multi-kernel one, then the
    # one-workgroup path (at most 8192 values), then an empty queue
    for n, hi in ((3_000_000, 1 << 26), (20_000, 1 << 30), (5_000, 1 << 20), (0, 1)):
        x = _queue(rng, n, hi)
        print("case noalloc n =", n, flush=True)
        got = _round_trip(L.g5_s4_roundtrip_noalloc, x)
        want = _round_trip(L.g5_s4_roundtrip, x)
        assert got[0] == want[0]
        assert np.array_equal(got[1], want[1])
        assert np.array_equal(got[2], x) and np.array_equal(want[2], x)
        if n <= 20_000:
            assert np.array_equal(got[1], O.s4_encode(x))


def test_noalloc_at_the_size_limit():
    g5.init(0)
    L = g5.lib()
    x = np.arange(MAX, dtype=np.uint32) * 3
    print("case noalloc n = MAX", flush=True)
    got = _round_trip(L.g5_s4_roundtrip_noalloc, x)
    want = _round_trip(L.g5_s4_roundtrip, x)
    assert got[0] == want[0] and np.array_equal(got[1], want[1]) and np.array_equal(got[2], x)


def test_noalloc_rejects_more_than_the_limit():
    g5.init(0)
    L = g5.lib()
    nw = g5.c_sz()
    one = g5.DeviceArray(16, np.uint32)
    # the size check comes before any device work: the buffers are not read
    rc = L.g5_s4_roundtrip_noalloc(one.ptr, MAX + 1, one.ptr, 16, one.ptr, ctypes.byref(nw))
    assert rc == -1 and b"at most" in L.g5_last_error()
