"""The CpuSimd gate on the GPU (cpu_2d/compression/cpusimd.hh:71-173): the
HipSimd schema of the C++ host side, driven through its Compression
interface by tests/hipsimd_gpu.cpp (built by gpugraph500_amd.build), against
the oracle's restatement of the reference's compress/decompress.

- above the threshold: the coded words are the reference codec's words
  (the oracle's orc_cpusimd_compress, pinned by the compiled reference
  codec's golden streams), and the receiver's rule reads them as coded;
- at or below the threshold: raw uint32 words;
- the H2 case (a coded stream exactly n words long, which the reference's
  receiver would mis-read as raw, cpusimd.hh:170-173): HipSimd sends the
  queue raw instead, and the receiver gets the queue back.
Every queue must come back identical after the receiver's decompress."""
import os
import subprocess

import numpy as np
import pytest

from gpugraph500_amd import build as gbuild
import oracle_lib as O

pytestmark = pytest.mark.gpu


def h2_queue():
    """260 values whose s4-bp128-d4 stream is 260 words long (the oracle
    test test_cpusimd_h2_rule_reproduced derives the count)."""
    k = np.arange(256)
    x = ((k // 4) % 4) * (1 << 30)
    tail = np.array([1 << 22, 2 << 22, 3 << 22, 4 << 22])
    return np.concatenate([x, tail]).astype(np.int64)


def run_schema(tmp_path, queues, threshold):
    exe = gbuild.HIPSIMD_TEST
    assert os.path.exists(exe), f"{exe} is missing: run python -m gpugraph500_amd.build"
    src, dst = str(tmp_path / "q.bin"), str(tmp_path / "w.bin")
    with open(src, "wb") as f:
        f.write(np.int32(threshold).tobytes())
        f.write(np.int64(len(queues)).tobytes())
        for q in queues:
            f.write(np.int64(len(q)).tobytes())
            f.write(np.ascontiguousarray(q, np.int64).tobytes())
    print("case hipsimd_gpu", threshold, [len(q) for q in queues], flush=True)
    r = subprocess.run([exe, src, dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    out = []
    with open(dst, "rb") as f:
        for _ in queues:
            cs = int(np.frombuffer(f.read(8), np.int64)[0])
            coded = int(np.frombuffer(f.read(4), np.int32)[0])
            out.append((coded, np.frombuffer(f.read(4 * cs), np.uint32).copy()))
    return out


def test_gate_above_at_and_below_threshold(tmp_path):
    rng = np.random.default_rng(8)
    thr = 64
    qs = []
    for n in (1, 63, 64, 65, 127, 128, 129, 1000, 4096 + 77, 200000):
        hi = max(4 * n, 1 << 12)
        qs.append(np.sort(rng.choice(hi, size=n, replace=False)).astype(np.int64))
    # a BFS level queue of the scale-16 Kronecker graph (sorted unique ids)
    o = O.build_csr(O.kronecker_edges(16, 16))
    root = int(O.select_roots(o, 1 << 16)[0])
    _, lev, nl = O.bfs_bin(o, root)
    qs += [np.flatnonzero(lev == d).astype(np.int64) for d in range(1, nl)]
    got = run_schema(tmp_path, qs, thr)
    for q, (coded, words) in zip(qs, got):
        want = O.cpusimd_compress(q, thr)
        if len(q) > thr and len(want) == len(q):    # an H2 stream: sent raw
            want = q.astype(np.uint32)
        assert np.array_equal(words, want), f"n = {len(q)}"
        assert coded == (len(q) > thr and len(want) != len(q))
        if len(q) <= thr:
            assert np.array_equal(words, q.astype(np.uint32))
        elif coded:
            assert np.array_equal(words, O.s4_encode(q.astype(np.uint32)))


def test_h2_stream_of_n_words_goes_raw(tmp_path):
    q = h2_queue()
    ref = O.cpusimd_compress(q, 64)
    assert len(ref) == len(q)                       # the reference's coded stream
    (coded, words), = run_schema(tmp_path, [q], 64)
    assert coded == 0                               # the receiver reads it as raw ...
    assert np.array_equal(words, q.astype(np.uint32))  # ... and raw is what HipSimd sent
    assert not np.array_equal(words, ref)


@pytest.mark.parametrize("thr", [1, 8, 200])
def test_thresholds(tmp_path, thr):
    rng = np.random.default_rng(thr)
    qs = [np.sort(rng.choice(1 << 20, size=n, replace=False)).astype(np.int64) for n in (thr, thr + 1, 3 * thr + 5)]
    for q, (coded, words) in zip(qs, run_schema(tmp_path, qs, thr)):
        assert np.array_equal(words, O.cpusimd_compress(q, thr))
        assert coded == (len(q) > thr)
