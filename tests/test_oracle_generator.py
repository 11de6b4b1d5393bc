"""Oracle generator and CSR build pinned by the reference generator's own
output (golden edge lists written by cpu_2d/generator compiled from source)
and by a brute-force CSR construction."""
import hashlib
import json

import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("scale", [5, 6, 7, 8])
def test_kronecker_matches_reference_edges(scale):
    ref = np.load(O.golden(f"kron_s{scale}.npy"))
    got = O.kronecker_edges(scale)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("scale", [10, 12, 16])
def test_kronecker_matches_reference_hash(scale):
    with open(O.golden("kron_hashes.json")) as f:
        h = json.load(f)[str(scale)]
    e = O.kronecker_edges(scale).reshape(-1)
    assert len(e) // 2 == h["m"]
    assert hashlib.sha256(e.tobytes()).hexdigest() == h["sha256"]


def test_kronecker_range_is_counter_based():
    # any sub-range equals the same slice of the full list (make_graph.c:25-31 splits by rank)
    full = O.kronecker_edges(8)
    part = O.kronecker_edges(8, start=1000, end=1777)
    assert np.array_equal(part, full[1000:1777])


def brute_csr(edges, n_rows):
    adj = [set() for _ in range(n_rows)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    rowptr = np.zeros(n_rows + 1, np.int64)
    cols = []
    for i in range(n_rows):
        s = sorted(adj[i])
        cols.extend(s)
        rowptr[i + 1] = rowptr[i] + len(s)
    return rowptr, np.array(cols, np.uint32)


@pytest.mark.parametrize("scale", [5, 6, 8])
def test_csr_matches_brute_force(scale):
    e = np.load(O.golden(f"kron_s{scale}.npy"))
    g = O.build_csr(e)
    assert g.n_rows % 64 == 0 and g.n_rows >= g.nglobal
    rp, cols = brute_csr(e.tolist(), g.n_rows)
    assert np.array_equal(g.rowptr, rp)
    assert np.array_equal(g.cols, cols)
    # generated directed degree counts self-loops twice and keeps duplicates
    gd = np.zeros(g.n_rows, np.int64)
    np.add.at(gd, e[:, 0], 1)
    np.add.at(gd, e[:, 1], 1)
    assert np.array_equal(g.gdeg, gd)


def test_csr_invariants_scale12():
    e = O.kronecker_edges(12)
    g = O.build_csr(e)
    rp, c = g.rowptr, g.cols.astype(np.int64)
    deg = np.diff(rp)
    rows = np.repeat(np.arange(g.n_rows), deg)
    assert (c != rows).all()                           # no self-loops
    same_row = rows[1:] == rows[:-1]
    assert (c[1:][same_row] > c[:-1][same_row]).all()  # strictly ascending per row
    key = rows * g.n_rows + c                          # symmetric
    rkey = c * g.n_rows + rows
    assert np.array_equal(np.sort(key), np.sort(rkey))


def test_padding_and_vertex_count():
    e = np.array([[0, 5], [5, 70]], np.int64)
    g = O.build_csr(e)
    assert g.nglobal == 71 and g.n_rows == 128
    assert g.nnz == 4
