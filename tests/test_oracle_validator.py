"""Mutation tests for the validator restatement (validate/validate.h:249-491):
a wrong parent, a level skip, a missing vertex and an extra vertex must each fail."""
import numpy as np

import oracle_lib as O


def setup():
    e = O.kronecker_edges(8)
    g = O.build_csr(e)
    root = int(O.select_roots(g, 256)[0])
    pred, lev, _ = O.bfs_bin(g, root)
    assert O.validate(e, g, 256, root, pred)[0] == 0
    return e, g, root, pred, lev


def test_wrong_parent_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    v = int(np.nonzero(lev == 2)[0][0])
    # a vertex at the same level is not adjacent-as-parent in general; pick a non-neighbour
    nbrs = set(g.cols[g.rowptr[v]:g.rowptr[v + 1]].tolist())
    cand = [u for u in np.nonzero(lev == 1)[0].tolist() if u not in nbrs]
    p[v] = cand[0]
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_level_skip_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    v = int(np.nonzero(lev == 3)[0][0])
    p[v] = root  # claims depth 1, but it is not a neighbour of the root
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_missing_vertex_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    v = int(np.nonzero(lev == 2)[0][0])
    p[v] = -1
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_extra_vertex_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    iso = [i for i in range(256) if g.rowptr[i + 1] == g.rowptr[i] and p[i] == -1]
    assert iso, "scale-8 graph has isolated vertices"
    p[iso[0]] = root
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_root_not_self_parent_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    p[root] = -1
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_self_parent_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    v = int(np.nonzero(lev == 1)[0][0])
    p[v] = v
    assert O.validate(e, g, 256, root, p)[0] != 0


def test_out_of_range_fails():
    e, g, root, pred, lev = setup()
    p = pred.copy()
    p[3] = 10 ** 9
    assert O.validate(e, g, 256, root, p)[0] != 0
