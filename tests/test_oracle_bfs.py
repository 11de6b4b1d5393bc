"""Oracle BFS (CPUBFS_bin restatement) pinned by brute force on tiny graphs,
closed forms, the validator restatement and SimpleCPUBFS's levels."""
import itertools
from collections import deque

import numpy as np
import pytest

import oracle_lib as O


def brute_bfs(n_rows, edges, root):
    adj = [set() for _ in range(n_rows)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    dist = [-1] * n_rows
    dist[root] = 0
    dq = deque([root])
    while dq:
        u = dq.popleft()
        for v in adj[u]:
            if dist[v] < 0:
                dist[v] = dist[u] + 1
                dq.append(v)
    pred = [-1] * n_rows
    pred[root] = root
    for v in range(n_rows):
        if dist[v] > 0:
            pred[v] = min(u for u in adj[v] if dist[u] == dist[v] - 1)
    return np.array(pred, np.int64), np.array(dist, np.int32)


def check_all_roots(edges, n_rows=None):
    edges = np.asarray(edges, np.int64).reshape(-1, 2)
    g = O.build_csr(edges, n_rows)
    for root in range(g.nglobal):
        if g.rowptr[root + 1] == g.rowptr[root]:
            continue
        pred, lev, nl = O.bfs_bin(g, root)
        bp, bl = brute_bfs(g.n_rows, edges.tolist(), root)
        assert np.array_equal(lev, bl), (root, lev, bl)
        assert np.array_equal(pred, bp), (root, pred, bp)
        assert nl == bl.max() + 1
        _, slev, snl = O.bfs_simple(g, root)
        assert np.array_equal(slev, lev) and snl == nl
        bad, depth, _ = O.validate(edges, g, g.n_rows, root, pred)
        assert bad == 0
        assert np.array_equal(depth, lev)


def test_random_tiny_graphs_all_roots():
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(2, 13))
        m = int(rng.integers(1, 3 * n))
        e = rng.integers(0, n, size=(m, 2))
        check_all_roots(e, 64)


def test_path():
    n = 9
    e = [(i, i + 1) for i in range(n - 1)]
    g = O.build_csr(np.array(e))
    pred, lev, nl = O.bfs_bin(g, 0)
    assert list(lev[:n]) == list(range(n)) and nl == n
    assert list(pred[:n]) == [0] + list(range(n - 1))
    pred, lev, nl = O.bfs_bin(g, 4)
    assert list(lev[:n]) == [4, 3, 2, 1, 0, 1, 2, 3, 4]


def test_star_and_complete():
    e = [(0, i) for i in range(1, 20)]
    g = O.build_csr(np.array(e))
    pred, lev, _ = O.bfs_bin(g, 7)
    assert lev[0] == 1 and all(lev[i] == 2 for i in range(1, 20) if i != 7)
    assert all(pred[i] == 0 for i in range(1, 20) if i != 7)
    e = list(itertools.combinations(range(10), 2))
    g = O.build_csr(np.array(e))
    pred, lev, nl = O.bfs_bin(g, 3)
    assert nl == 2 and all(pred[i] == 3 for i in range(10))


def test_disconnected_components():
    e = [(0, 1), (1, 2), (10, 11)]
    g = O.build_csr(np.array(e))
    pred, lev, _ = O.bfs_bin(g, 10)
    assert lev[11] == 1 and pred[11] == 10
    assert (lev[[0, 1, 2]] == -1).all() and (pred[[0, 1, 2]] == -1).all()


def test_duplicates_and_self_loops_do_not_change_result():
    e = [(0, 1), (1, 0), (0, 1), (2, 2), (1, 2), (2, 1)]
    g = O.build_csr(np.array(e))
    pred, lev, _ = O.bfs_bin(g, 0)
    assert list(lev[:3]) == [0, 1, 2] and list(pred[:3]) == [0, 0, 1]


@pytest.mark.parametrize("scale", [8, 10])
def test_kronecker_all_roots_valid(scale):
    e = O.kronecker_edges(scale)
    g = O.build_csr(e)
    roots = O.select_roots(g, 1 << scale)
    assert len(roots) == 64 and len(set(roots.tolist())) == 64
    for r in roots:
        pred, lev, nl = O.bfs_bin(g, int(r))
        bad, depth, ne = O.validate(e, g, 1 << scale, int(r), pred)
        assert bad == 0
        assert np.array_equal(depth, lev)
        assert ne == O.nedge(g, lev)
        _, slev, _ = O.bfs_simple(g, int(r))
        assert np.array_equal(slev, lev)
    # brute force on one root as well
    bp, bl = brute_bfs(g.n_rows, e.tolist(), int(roots[0]))
    pred, lev, _ = O.bfs_bin(g, int(roots[0]))
    assert np.array_equal(pred, bp) and np.array_equal(lev, bl)


def test_simple_bfs_parents_are_max_id():
    e = [(0, 1), (0, 2), (1, 3), (2, 3)]
    g = O.build_csr(np.array(e))
    pb, _, _ = O.bfs_bin(g, 0)
    ps, _, _ = O.bfs_simple(g, 0)
    assert pb[3] == 1 and ps[3] == 2
