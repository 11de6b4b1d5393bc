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


RANGE, ROOT, SELF, DEPTH, EDGE, PARENT = 1, 2, 4, 8, 16, 32


def test_failure_flags_name_the_check():
    """Each mutation raises the flag of the check that catches it
    (validate.h:258-264 range, :285-295 root, :300-342 self parent,
    :425-433 edge levels, :436-471 parent edges)."""
    e, g, root, pred, lev = setup()
    p = pred.copy()
    p[3] = 10 ** 9
    assert O.validate(e, g, 256, root, p)[0] == RANGE
    assert O.validate(e, g, 256, 256, pred)[0] == RANGE
    p = pred.copy()
    p[root] = -1
    assert O.validate(e, g, 256, root, p)[0] & ROOT
    p = pred.copy()
    v = int(np.nonzero(lev == 1)[0][0])
    p[v] = v
    assert O.validate(e, g, 256, root, p)[0] & SELF
    p = pred.copy()
    v = int(np.nonzero(lev == 2)[0][0])
    p[v] = -1  # v unreached, its neighbours reached
    assert O.validate(e, g, 256, root, p)[0] & EDGE


def path_edges(n):
    return np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int64)


def test_depth_saturates_at_65535():
    """Depths are 16-bit with 65535 meaning unreached: on a path longer than
    that, vertices from distance 65535 on are unreached, and every check
    still passes (build_bfs_depth_map, validate.h:110-242)."""
    n = 70000
    e = path_edges(n)
    g = O.build_csr(e)
    pred = np.full(g.n_rows, -1, np.int64)
    pred[0] = 0
    pred[1:n] = np.arange(n - 1)
    bad, depth, ne = O.validate(e, g, n, 0, pred)
    assert bad == 0
    assert depth[65534] == 65534 and depth[65535] == -1 and depth[n - 1] == -1
    assert ne == n - 1


def test_parent_cycle_off_the_root_component_passes():
    """The reference accepts a parent cycle inside a component the root does
    not reach: its vertices keep depth 65535 and their parent edges exist."""
    e = np.array([[0, 1], [1, 2], [3, 4]], np.int64)
    g = O.build_csr(e)
    pred = np.full(g.n_rows, -1, np.int64)
    pred[0], pred[1], pred[2] = 0, 0, 1
    assert O.validate(e, g, 5, 0, pred)[0] == 0
    pred[3], pred[4] = 4, 3
    bad, depth, ne = O.validate(e, g, 5, 0, pred)
    assert bad == 0 and depth[3] == -1 and depth[4] == -1 and ne == 3
