"""Single-root BFS on the device: levels and minimum-id parents bit-exact
against the oracle's CPUBFS_bin restatement for every root, the hook path
(start/expand/outqueue/set_outqueue/set_inqueue) equal to the resident loop,
and every result accepted by the validator restatement."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O
from test_gpu_continuous import checksum

pytestmark = pytest.mark.gpu

_CACHE = {}


def kron(scale):
    if scale not in _CACHE:
        e = O.kronecker_edges(scale)
        ref = O.build_csr(e)
        g = g5.Graph.upload(ref.rowptr, ref.cols, ref.n_rows, gdeg=ref.gdeg)
        _CACHE[scale] = (e, ref, g)
    return _CACHE[scale]


@pytest.mark.parametrize("scale", [10, 16])
def test_resident_all_roots_bit_exact(scale):
    e, ref, g = kron(scale)
    roots = O.select_roots(ref, 1 << scale)
    assert len(roots) == 64
    b = g5.BFS(g)
    for r in roots.tolist():
        nl, ms = b.run(r)
        pred, lev = b.result()
        opred, olev, onl = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev), f"levels differ, root {r}"
        assert np.array_equal(pred, opred), f"parents differ, root {r}"
        assert nl == onl
        assert b.nedge() == O.nedge(ref, olev)
        assert b.summary() == (O.nedge(ref, olev), checksum(opred))
    bad, depth, ne = O.validate(e, ref, 1 << scale, int(roots[-1]), pred)
    assert bad == 0 and np.array_equal(depth, lev)


@pytest.mark.parametrize("scale", [14, 18])
def test_directions_identical(scale):
    """Direction-optimised (bottom-up dense levels) and top-down-only searches
    give the same levels and parents, equal to the oracle's; the bottom-up
    kernel really ran."""
    e, ref, g = kron(scale)
    roots = O.select_roots(ref, 1 << scale)[:16]
    b = g5.BFS(g)
    b.profile(True)
    for r in roots.tolist():
        nl_a, _ = b.run(r)
        pa, la = b.result()
        nl_b, _ = b.run(r, topdown=True)
        pb, lb = b.result()
        opred, olev, onl = O.bfs_bin(ref, r)
        assert nl_a == nl_b == onl
        assert np.array_equal(la, olev) and np.array_equal(lb, olev), f"levels differ, root {r}"
        assert np.array_equal(pa, opred) and np.array_equal(pb, opred), f"parents differ, root {r}"
    ks = b.kernel_stats()
    assert ks["k_bottom_up"]["launches"] > 0 and ks["k_expand"]["launches"] > 0


@pytest.mark.parametrize("alpha,beta", [(1, 1), (1, 1 << 30), (1 << 30, 1), (2, 4), (64, 64)])
def test_direction_switch_parameters_do_not_change_results(alpha, beta):
    # alpha 1 / beta 2^30: bottom-up from the second level on, never back;
    # alpha 2^30: top-down throughout; the others switch at other levels
    e, ref, g = kron(13)
    roots = O.select_roots(ref, 1 << 13)[:8]
    b = g5.BFS(g)
    b.tune(alpha, beta)
    for r in roots.tolist():
        nl, _ = b.run(r)
        pred, lev = b.result()
        opred, olev, onl = O.bfs_bin(ref, r)
        assert nl == onl and np.array_equal(lev, olev) and np.array_equal(pred, opred), (alpha, beta, r)
    with pytest.raises(g5.G5Error):
        b.tune(0, 24)


def test_hook_path_matches_oracle_levels_and_queues():
    e, ref, g = kron(12)
    roots = O.select_roots(ref, 1 << 12)[:6]
    b = g5.BFS(g)
    for r in roots.tolist():
        opred, olev, onl = O.bfs_bin(ref, r)
        b.start(r)
        depth = 0
        while True:
            any_new = b.expand()
            if depth > 0 and not any_new:
                break
            q = b.outqueue()
            assert np.array_equal(q, np.nonzero(olev == depth + 1)[0].astype(np.uint32))
            b.set_outqueue(q if depth % 2 else None)   # reduced queue from "vreduce", or self
            b.set_inqueue(q if depth % 2 == 0 else None)  # row queue from the fold, or own
            depth += 1
        pred, lev = b.result()
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred)
        assert depth + 1 == onl


def test_codec_in_the_loop_gives_same_result():
    e, ref, g = kron(16)
    roots = O.select_roots(ref, 1 << 16)[:8]
    b = g5.BFS(g)
    for r in roots.tolist():
        b.run(r, codec=True)
        pred, lev = b.result()
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred)


def test_tiny_graphs_all_roots():
    rng = np.random.default_rng(123)
    for trial in range(25):
        n = int(rng.integers(2, 40))
        m = int(rng.integers(1, 4 * n))
        edges = rng.integers(0, n, size=(m, 2))
        ref = O.build_csr(edges, 64)
        g = g5.Graph.upload(ref.rowptr, ref.cols, 64, gdeg=ref.gdeg)
        b = g5.BFS(g)
        for r in range(n):
            if ref.rowptr[r + 1] == ref.rowptr[r]:
                continue
            b.run(r)
            pred, lev = b.result()
            opred, olev, _ = O.bfs_bin(ref, r)
            assert np.array_equal(lev, olev) and np.array_equal(pred, opred), (trial, r)
        b.free()
        g.free()


def test_isolated_root_and_bad_root():
    ref = O.build_csr(np.array([[0, 1], [1, 2], [5, 5]]), 64)
    g = g5.Graph.upload(ref.rowptr, ref.cols, 64, gdeg=ref.gdeg)
    b = g5.BFS(g)
    nl, _ = b.run(5)
    pred, lev = b.result()
    assert nl == 1 and pred[5] == 5 and lev[5] == 0 and (lev[[0, 1, 2]] == -1).all()
    with pytest.raises(g5.G5Error):
        b.run(64)
    with pytest.raises(g5.G5Error):
        b.set_inqueue(np.array([3, 2], np.uint32))  # not sorted


def test_outqueue_capacity_error():
    e, ref, g = kron(10)
    b = g5.BFS(g)
    b.start(int(O.select_roots(ref, 1024)[0]))
    b.expand()
    import ctypes
    q = np.empty(1, np.uint32)
    n = ctypes.c_int64(0)
    rc = g5.lib().g5_bfs_outqueue(b.h, q.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n))
    assert rc == -4  # G5_E_OVERFLOW
