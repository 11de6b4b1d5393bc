"""Serving (g500_serve): several resident graphs, a Poisson query stream,
continuous batching into the 64-lane search.  Every query's result must be
the same served in the mix as run alone, and equal to the oracle's search."""
import json
import subprocess

import numpy as np
import pytest

from gpugraph500_amd import build as gbuild
import oracle_lib as O
from test_gpu_continuous import checksum

pytestmark = pytest.mark.gpu


def _serve(tmp_path, *args):
    dump = str(tmp_path / "serve.txt")
    cmd = [gbuild.SERVE, *args, "-dump", dump]
    print("case", cmd, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    rows = []
    with open(dump) as f:
        for line in f:
            t = line.split()
            qid, scale, arr, done, n = int(t[0]), int(t[1]), float(t[2]), float(t[3]), int(t[4])
            roots = [(int(t[5 + 3 * k]), int(t[6 + 3 * k]), int(t[7 + 3 * k])) for k in range(n)]
            rows.append((qid, scale, arr, done, roots))
    return rec, rows


def test_served_equals_alone_under_overlap(tmp_path):
    """A burst far above the service rate: queries overlap on every engine
    and join running searches between levels."""
    rec, rows = _serve(tmp_path, "-g", "10,12,14", "-rate", "50000", "-nq", "300", "-seed", "3", "-compare")
    assert rec["mismatches"] == 0 and rec["queries"] == 300
    assert all(done >= arr for _, _, arr, done, _ in rows)
    # the engines really did run several queries' roots at once
    assert rec["levels"] < sum(len(r[4]) for r in rows) * 4


def test_served_equals_alone_with_few_lanes(tmp_path):
    """Eight lanes per engine: admission waits for lanes to free up."""
    rec, rows = _serve(tmp_path, "-g", "11,13", "-rate", "20000", "-nq", "200", "-seed", "5", "-lanes", "8",
                       "-compare")
    assert rec["mismatches"] == 0


def test_served_results_match_oracle(tmp_path):
    rec, rows = _serve(tmp_path, "-g", "12", "-rate", "20000", "-nq", "40", "-seed", "7")
    o = O.build_csr(O.kronecker_edges(12, 16))
    seen = {}
    for _, scale, _, _, roots in rows:
        for root, ne, ck in roots:
            if root in seen:
                assert seen[root] == (ne, ck)
                continue
            rp, rl, _ = O.bfs_bin(o, root)
            assert ne == O.nedge(o, rl), root
            assert ck == checksum(rp), root
            seen[root] = (ne, ck)


def test_alone_mode(tmp_path):
    rec, rows = _serve(tmp_path, "-g", "10", "-nq", "20", "-alone")
    assert rec["mode"] == "alone" and len(rows) == 20


@pytest.mark.parametrize("small", ["2", "8"])
def test_small_queries_on_idle_engines(tmp_path, small):
    """A slow stream: most queries find their engine idle, and those of at
    most `small` roots are answered by single-root searches -- with the same
    (nedge, checksum) as the multi-lane search run alone, and the oracle's."""
    rec, rows = _serve(tmp_path, "-g", "12", "-rate", "200", "-nq", "30", "-seed", "4", "-small", small,
                       "-compare")
    assert rec["mismatches"] == 0 and rec["small"] == int(small)
    assert rec["single_roots"] > 0
    o = O.build_csr(O.kronecker_edges(12, 16))
    for _, _, _, _, roots in rows:
        for root, ne, ck in roots:
            rp, rl, _ = O.bfs_bin(o, root)
            assert ne == O.nedge(o, rl) and ck == checksum(rp), root


@pytest.mark.parametrize("window", ["-1", "0", "3"])
def test_join_window_does_not_change_results(tmp_path, window):
    """Joining whenever lanes are free (-1), only into an idle engine (0),
    or within a cohort's first 3 levels: the same results."""
    rec, rows = _serve(tmp_path, "-g", "11,13", "-rate", "50000", "-nq", "150", "-seed", "9", "-window", window,
                       "-compare")
    assert rec["mismatches"] == 0 and rec["window"] == int(window)


def test_record_placement_at_start_does_not_change_results(tmp_path):
    """Engines that pick their parent records' placement at start (-place)
    serve the same results as alone."""
    rec, rows = _serve(tmp_path, "-g", "11,13", "-rate", "20000", "-nq", "120", "-seed", "4", "-place", "2",
                       "-compare")
    assert rec["mismatches"] == 0 and rec["place"] == 2


def test_memory_budget_holds_queries_back(tmp_path):
    """Admission by memory: the running roots may hold 3.5 lanes' bytes of the
    scale-12 search at once.  Queries of up to 3 roots arrive in a burst;
    those whose roots' bytes are not free wait, and every query completes
    with the result it has alone."""
    import gpugraph500_amd as g5
    g = g5.Graph.kronecker(12)
    _, _, lane = g5.MSBFS.footprint(g)
    g.free()
    budget = 3.5 * lane
    rec, rows = _serve(tmp_path, "-g", "12", "-rate", "50000", "-nq", "120", "-seed", "9", "-maxroots", "3",
                       "-small", "0", "-budget", repr(budget / 1048576.0), "-compare")
    assert rec["mismatches"] == 0 and rec["queries"] == 120
    assert rec["engines"]["12"]["lane_bytes"] == lane
    assert rec["budget_peak_bytes"] <= rec["budget_bytes"] < 4 * lane
    assert rec["budget_peak_bytes"] >= lane
    assert rec["budget_waits"] > 0
    assert all(done >= arr for _, _, arr, done, _ in rows)


def test_memory_budget_rejects_what_can_never_fit(tmp_path):
    """A query whose roots need more than the whole budget is an error, and a
    graph that does not fit the device's free memory is not admitted."""
    import gpugraph500_amd as g5
    g = g5.Graph.kronecker(10)
    _, _, lane = g5.MSBFS.footprint(g)
    g.free()
    cmd = [gbuild.SERVE, "-g", "10", "-nq", "20", "-maxroots", "4", "-small", "0", "-seed", "2",
           "-budget", repr(1.5 * lane / 1048576.0)]
    print("case", cmd, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "whole budget" in r.stderr
    # scale 33: about 2.3 TB of search state, checked before anything is built
    r = subprocess.run([gbuild.SERVE, "-g", "33", "-nq", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "not admitted" in r.stderr
