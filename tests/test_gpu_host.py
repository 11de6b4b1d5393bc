"""The C++ host side on the GPU: the g500_hip driver's emulated R x C grid
(HipBFS hooks, vreduce, fold, codec on the wire) against the oracle, the
2D block builder against the oracle CSR, and the device bitmap/queue ABI."""
import os
import subprocess

import numpy as np
import pytest

import gpugraph500_amd as g5
from gpugraph500_amd import build as gbuild
import oracle_lib as O

pytestmark = pytest.mark.gpu

SCALE = 12


@pytest.fixture(scope="module")
def oracle_csr():
    return O.build_csr(O.kronecker_edges(SCALE, 16))


def _read_dump(path):
    out = []
    with open(path, "rb") as f:
        k = int(np.frombuffer(f.read(8), np.int64)[0])
        for _ in range(k):
            root, n = np.frombuffer(f.read(16), np.int64)
            pred = np.frombuffer(f.read(8 * int(n)), np.int64).copy()
            out.append((int(root), pred))
    return out


@pytest.mark.parametrize("R,x,codec", [
    (1, "device", ""),
    (2, "host", ""),
    (2, "host", "s4-bp128-d4"),
    (2, "device", "s4-bp128-d4"),
    (3, "device", ""),
    (3, "host", "s4-bp128-d4"),
    (4, "device", "s4-bp128-d4"),
    (4, "host", "s4-bp128-d4"),
    (2, "device-td", ""),
    (4, "device-td", "s4-bp128-d4"),
])
def test_grid_matches_oracle(tmp_path, oracle_csr, R, x, codec):
    """Every search of the emulated grid equals the 1 x 1 search (-verify) and
    the dumped predecessors are the oracle's minimum-id parents; the device
    exchange runs its dense levels bottom-up unless -td."""
    dump = str(tmp_path / "pred.bin")
    cmd = [gbuild.DRIVER, "-s", str(SCALE), "-R", str(R), "-C", str(R), "-n", "6", "-x", x.split("-")[0], "-btr",
           "16", "-verify", "-dump", dump, "-dumpk", "3"] + (["-td"] if x.endswith("-td") else [])
    if codec:
        cmd += ["-be", codec]
    print("case", cmd, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "verify_mismatches: 0" in r.stdout
    if x == "device":
        import json
        assert json.loads(r.stdout.strip().splitlines()[-1])["bottom_up_levels"] > 0
    for root, pred in _read_dump(dump):
        ref, _, _ = O.bfs_bin(oracle_csr, root)
        n = min(len(ref), len(pred))
        assert np.array_equal(pred[:n], ref[:n]), f"root {root}"
        assert (pred[n:] == -1).all() and (ref[n:] == -1).all()


def test_grid_codec_shrinks_the_fold(tmp_path):
    """With the codec the fold's queues cross in fewer words than raw."""
    cmd = [gbuild.DRIVER, "-s", str(SCALE), "-R", "2", "-C", "2", "-n", "4", "-x", "host", "-be", "s4-bp128-d4",
           "-btr", "16"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import json
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert 0 < rec["coded_words"] < rec["raw_words"]


@pytest.mark.parametrize("R", [2, 3])
def test_kronecker_2d_blocks_reassemble(oracle_csr, R):
    """The R x R blocks, put back together with their bounds, give the oracle
    CSR row by row, and the blocks' generated degrees sum to the oracle's."""
    blocks = g5.Graph.kronecker_2d(SCALE, R, R)
    o = oracle_csr
    n = o.n_rows
    rows = [[] for _ in range(R)]
    gsum = np.zeros(n + 64 * R, np.int64)
    for i, b in enumerate(blocks):
        r, c = divmod(i, R)
        rs, nr, cs, nc = g5.block_bounds(b.nglobal, R, R, r, c)
        assert (b.row_start, b.n_rows, b.col_start, b.n_cols) == (rs, nr, cs, nc)
        rp, cols, gdeg = b.download()
        assert (cols < nc).all()
        rows[r].append((cs, rp, cols))
        gsum[rs:rs + nr] += gdeg
    for r in range(R):
        rs = blocks[r * R].row_start
        for i in range(blocks[r * R].n_rows):
            v = rs + i
            got = np.concatenate([cs + cols[rp[i]:rp[i + 1]].astype(np.int64) for cs, rp, cols in rows[r]])
            want = o.cols[o.rowptr[v]:o.rowptr[v + 1]].astype(np.int64) if v < n else np.empty(0, np.int64)
            assert np.array_equal(got, want), f"row {v}"
    assert np.array_equal(gsum[:n], o.gdeg[:n])
    assert (gsum[n:] == 0).all()


def test_bitmap_queue_abi():
    rng = np.random.default_rng(5)
    words = 1000
    ids = np.unique(rng.integers(0, words * 64, 20000)).astype(np.uint32)
    dq = g5.DeviceArray.from_host(ids)
    bits = g5.DeviceArray(words, np.uint64)
    g5.check(g5.lib().g5_queue_to_bitmap(dq.ptr, len(ids), bits.ptr, words))
    hb = bits.to_host()
    ref = np.zeros(words, np.uint64)
    np.bitwise_or.at(ref, ids >> 6, np.left_shift(np.uint64(1), (ids & 63).astype(np.uint64)))
    assert np.array_equal(hb, ref)
    out = g5.DeviceArray(words * 64, np.uint32)
    n = g5.c_i64()
    g5.check(g5.lib().g5_bitmap_to_queue(bits.ptr, words, out.ptr, g5.ctypes.byref(n)))
    assert n.value == len(ids)
    assert np.array_equal(out.to_host(n.value), ids)
    # an id beyond the bitmap is rejected, nothing is written out of range
    bad = g5.DeviceArray.from_host(np.array([5, words * 64 + 3], np.uint32))
    rc = g5.lib().g5_queue_to_bitmap(bad.ptr, 2, bits.ptr, words)
    assert rc == -1 and b"beyond" in g5.lib().g5_last_error()
