"""Parity at the benchmark configurations' own sizes, as the reference checks
every root of its run at the run's scale (cpu_2d/main.cpp:500-519 ->
validate/validate.h:249-491):

- single-root search with the s4-bp128-d4 codec on every level's queue at
  scale 22 (levels and minimum-id parents against the oracle's CPUBFS_bin
  restatement; the codec's word streams against the oracle encoder, which
  the compiled reference codec's golden streams pin);
- the batched 64-root search at scales 24 and 26: all 64 lanes' predecessor
  and level maps equal, on the device, to the single-root search of the same
  root, every lane accepted by the device validator with the same TEPS
  numerator, and one lane against the oracle;
- one rank's share of a 4 x 4 grid at scale 26 (the g500_hip driver: every
  search equal to the 1 x 1 search, one root against the oracle);
- serving over the six resident graphs (scales 18-26) at two arrival rates,
  every query's result the same served in the mix as run alone.
"""
import json
import subprocess
import time

import numpy as np
import pytest

import gpugraph500_amd as g5
from gpugraph500_amd import build as gbuild
import oracle_lib as O

pytestmark = pytest.mark.gpu


def host_csr(g):
    """The device-built graph as an oracle CSR (the device CSR equals the
    oracle's own build: tests/test_gpu_graph.py)."""
    rp, cols, gdeg = g.download()
    return O.CSR(rp, cols, g.n_rows, gdeg, g.nglobal)


def log(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------- C2

def test_c2_codec_in_the_loop_scale22():
    scale = 22
    t0 = time.time()
    g = g5.Graph.kronecker(scale)
    roots = g.select_roots(1 << scale, 64)
    ref = host_csr(g)
    b = g5.BFS(g)
    plain = g5.DeviceArray(g.n_rows, np.int64)
    coded = g5.DeviceArray(g.n_rows, np.int64)
    lplain = g5.DeviceArray(g.n_rows, np.int32)
    lcoded = g5.DeviceArray(g.n_rows, np.int32)
    # every root: the search with the codec on every level equals the plain one
    for r in roots.tolist():
        b.run(r)
        b.pred_device(plain)
        b.level_device(lplain)
        b.run(r, codec=True)
        b.pred_device(coded)
        b.level_device(lcoded)
        assert g5.diff_u32(plain, coded, 2 * g.n_rows) == 0, f"root {r}"
        assert g5.diff_u32(lplain, lcoded, g.n_rows) == 0, f"root {r}"
    log(f"c2: 64 roots, codec == plain ({time.time() - t0:.1f} s)")
    # 8 roots against the oracle, levels and parents
    for r in roots[:8].tolist():
        b.run(r, codec=True)
        pred, lev = b.result()
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev), f"levels, root {r}"
        assert np.array_equal(pred, opred), f"parents, root {r}"
    log(f"c2: 8 roots == oracle ({time.time() - t0:.1f} s)")
    # 2 roots: every level's queue, coded on the device by the same round
    # trip the search runs, word for word against the oracle encoder
    for r in roots[:2].tolist():
        _, olev, nl = O.bfs_bin(ref, r)
        for d in range(1, nl):
            q = np.flatnonzero(olev == d).astype(np.uint32)
            back, nw, words = g5.s4_roundtrip(q, words=True)
            assert np.array_equal(back, q)
            assert np.array_equal(words, O.s4_encode(q)), f"root {r} level {d} ({len(q)} ids)"
    log(f"c2: streams of 2 roots == oracle ({time.time() - t0:.1f} s)")
    b.free()
    g.free()


# ---------------------------------------------------------------- C4

@pytest.mark.parametrize("scale,shape", [
    (24, None),
    (26, None),
    (26, dict(thread_max=256, wave_max=2048, lane_ratio=256, split_max=32)),
])
def test_c4_batched_equals_single_root(scale, shape):
    t0 = time.time()
    g = g5.Graph.kronecker(scale)
    roots = g.select_roots(1 << scale, 64)
    n = g.n_rows
    m = g5.MSBFS(g)
    if shape:
        m.tune(**shape)
    nl, ms = m.run(roots)
    s_ms, f_ms = m.times()
    log(f"c4 s{scale} {shape}: batched {ms:.1f} ms (search {s_ms:.1f}, finalise {f_ms:.2f}), {nl} levels")
    _, _, lane_of, final = m.maps()
    assert final and sorted(lane_of.tolist()) == list(range(64))
    ne_b = m.nedge()
    b = g5.BFS(g)
    val = g5.Validator(g)
    ps, pb = g5.DeviceArray(n, np.int64), g5.DeviceArray(n, np.int64)
    ls, lb = g5.DeviceArray(n, np.int32), g5.DeviceArray(n, np.int32)
    for lane, r in enumerate(roots.tolist()):
        b.run(r)
        b.pred_device(ps)
        b.level_device(ls)
        m.pred_device(lane, pb)
        m.level_device(lane, lb)
        assert g5.diff_u32(ps, pb, 2 * n) == 0, f"parents differ, lane {lane} root {r}"
        assert g5.diff_u32(ls, lb, n) == 0, f"levels differ, lane {lane} root {r}"
        bad, vne, _, _ = val.validate_device(r, pb)
        assert bad == 0, f"validator flags {bad}, lane {lane} root {r}"
        assert vne == ne_b[lane] == b.nedge(), f"nedge, lane {lane} root {r}"
    log(f"c4 s{scale}: 64 lanes == single-root, validated ({time.time() - t0:.1f} s)")
    if shape is None:
        ref = host_csr(g)
        lane = 0 if scale == 26 else 63
        r = int(roots[lane])
        pred, lev = m.result(lane)
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred), f"lane {lane} root {r} vs oracle"
        log(f"c4 s{scale}: lane {lane} == oracle ({time.time() - t0:.1f} s)")
    for a in (ps, pb, ls, lb):
        a.free()
    val.free()
    b.free()
    m.free()
    g.free()


# ---------------------------------------------------------------- C3

def _read_dump(path):
    out = []
    with open(path, "rb") as f:
        k = int(np.frombuffer(f.read(8), np.int64)[0])
        for _ in range(k):
            root, n = np.frombuffer(f.read(16), np.int64)
            out.append((int(root), np.frombuffer(f.read(8 * int(n)), np.int64).copy()))
    return out


def test_c3_grid_4x4_scale26(tmp_path):
    dump = str(tmp_path / "pred.bin")
    cmd = [gbuild.DRIVER, "-s", "26", "-R", "4", "-C", "4", "-n", "4", "-x", "device", "-be", "s4-bp128-d4",
           "-share", "0", "0", "-verify", "-dump", dump, "-dumpk", "1"]
    log("case " + " ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "verify_mismatches: 0" in r.stdout
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert rec["roots"] == 4
    g = g5.Graph.kronecker(26)
    ref = host_csr(g)
    g.free()
    (root, pred), = _read_dump(dump)
    opred, _, _ = O.bfs_bin(ref, root)
    k = min(len(pred), len(opred))
    assert np.array_equal(pred[:k], opred[:k]), f"root {root}"
    assert (pred[k:] == -1).all() and (opred[k:] == -1).all()


# ---------------------------------------------------------------- C5

@pytest.mark.parametrize("rate,nq", [(200, 200), (5000, 300)])
def test_c5_six_graphs_served_equals_alone(tmp_path, rate, nq):
    dump = str(tmp_path / "serve.txt")
    cmd = [gbuild.SERVE, "-g", "18,20,22,24,25,26", "-rate", str(rate), "-nq", str(nq), "-seed", "11", "-compare",
           "-dump", dump]
    log("case " + " ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert rec["mismatches"] == 0 and rec["queries"] == nq
    log(f"c5 rate {rate}: p50 {rec.get('p50_ms')} ms, p99 {rec.get('p99_ms')} ms, {rec.get('gteps')} GTEPS")
    # the scale-18 graph's answers against the oracle's searches
    o = O.build_csr(O.kronecker_edges(18, 16))
    from test_gpu_continuous import checksum
    seen = 0
    with open(dump) as f:
        for line in f:
            t = line.split()
            if int(t[1]) != 18:
                continue
            for k in range(int(t[4])):
                root, ne, ck = int(t[5 + 3 * k]), int(t[6 + 3 * k]), int(t[7 + 3 * k])
                rp, rl, _ = O.bfs_bin(o, root)
                assert ne == O.nedge(o, rl) and ck == checksum(rp), f"root {root}"
                seen += 1
            if seen >= 16:
                break
    assert seen > 0
