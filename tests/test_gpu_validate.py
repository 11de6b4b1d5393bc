"""Device result validation (validate/validate.h:249-491): the failure flags,
the TEPS numerator and the depth map equal the oracle's restatement, which
checks against the generated edge list while the device checks the CSR --
on correct searches, on every mutation the reference's checks catch (and
the ones they let through), on rows split into pieces, and on depths past
the 16-bit range."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu

RANGE, ROOT, SELF, EDGE, PARENT = 1, 2, 4, 16, 32
_CACHE = {}


def uploaded(edges, key):
    if key not in _CACHE:
        ref = O.build_csr(edges)
        g = g5.Graph.upload(ref.rowptr, ref.cols, ref.n_rows, gdeg=ref.gdeg)
        _CACHE[key] = (edges, ref, g, g5.Validator(g))
    return _CACHE[key]


def kron(scale):
    return uploaded(O.kronecker_edges(scale), ("kron", scale))


def same_as_oracle(e, ref, val, root, pred, nglobal=None):
    nglobal = ref.nglobal if nglobal is None else nglobal
    obad, odepth, one = O.validate(e, ref, nglobal, root, pred)
    bad, ne, depth, ms = val.validate(root, pred, nglobal=nglobal, want_depth=True)
    assert bad == obad, (bad, obad)
    if not obad & RANGE:
        assert ne == one
        assert np.array_equal(depth, odepth)
    return bad


@pytest.mark.parametrize("scale", [10, 16])
def test_correct_searches_pass(scale):
    e, ref, g, val = kron(scale)
    roots = O.select_roots(ref, ref.nglobal)[:8]
    b = g5.BFS(g)
    for r in roots.tolist():
        b.run(r)
        pred, lev = b.result()
        bad, ne, depth, _ = val.validate(r, pred, want_depth=True)
        assert bad == 0, f"root {r}: flags {bad}"
        assert ne == b.nedge() == O.nedge(ref, lev)
        assert np.array_equal(depth, lev)
        assert same_as_oracle(e, ref, val, r, pred) == 0


def mutations(ref, root, pred, lev, rng):
    """(name, pred) pairs: each breaks the map in one way."""
    n = ref.n_rows
    out = []
    p = pred.copy(); p[int(rng.integers(n))] = 10 ** 9; out.append(("parent out of range", p))
    p = pred.copy(); p[int(rng.integers(n))] = -2; out.append(("parent below -1", p))
    p = pred.copy(); p[root] = -1; out.append(("root unreached", p))
    at = {d: np.nonzero(lev == d)[0] for d in range(1, 4)}
    v = int(at[1][0])
    p = pred.copy(); p[v] = v; out.append(("self parent", p))
    v = int(at[2][0])
    p = pred.copy(); p[v] = -1; out.append(("missing vertex", p))
    nbrs = set(ref.cols[ref.rowptr[v]:ref.rowptr[v + 1]].tolist())
    far = [u for u in at[1].tolist() if u not in nbrs]
    p = pred.copy(); p[v] = far[0]; out.append(("non-adjacent parent", p))
    if len(at[3]):
        w = int(at[3][0])
        p = pred.copy(); p[w] = root; out.append(("level skip", p))
    iso = np.nonzero((np.diff(ref.rowptr) == 0) & (pred == -1))[0]
    if len(iso):
        p = pred.copy(); p[int(iso[0])] = root; out.append(("extra vertex", p))
    a, b = int(at[2][0]), int(at[2][-1])
    p = pred.copy(); p[a], p[b] = b, a; out.append(("two-cycle", p))
    p = rng.integers(-1, n, size=n).astype(np.int64); p[root] = root; out.append(("random map", p))
    p = pred.copy(); p[lev == 3] = -1; out.append(("level 3 dropped", p))
    return out


@pytest.mark.parametrize("scale", [10, 14])
def test_mutations_flagged_as_the_oracle_does(scale):
    e, ref, g, val = kron(scale)
    rng = np.random.default_rng(scale)
    root = int(O.select_roots(ref, ref.nglobal)[0])
    pred, lev, _ = O.bfs_bin(ref, root)
    for name, p in mutations(ref, root, pred, lev, rng):
        bad = same_as_oracle(e, ref, val, root, p)
        assert bad != 0, name
    assert same_as_oracle(e, ref, val, ref.nglobal, pred) == RANGE  # root out of range


def test_heavy_rows_split_into_pieces():
    """A hub of 5000 neighbours (3 pieces) and a second hub reached at depth 2:
    the piece kernel and the parent binary search decide as the oracle does."""
    hub, leaves = 0, np.arange(1, 5001)
    e = [np.stack([np.full(5000, hub), leaves], axis=1)]
    hub2 = 5001
    e.append(np.array([[1, hub2]]))
    e.append(np.stack([np.full(3000, hub2), np.arange(5002, 8002)], axis=1))
    e = np.concatenate(e).astype(np.int64)
    e, ref, g, val = uploaded(e, "hubs")
    for root in (hub, hub2, 7000):
        pred, lev, _ = O.bfs_bin(ref, root)
        assert same_as_oracle(e, ref, val, root, pred) == 0
        p = pred.copy(); p[4000] = 17 if root != 17 else 18  # leaf claims a leaf
        assert same_as_oracle(e, ref, val, root, p) & PARENT
        p = pred.copy(); p[6000] = -1  # a leaf of the second hub dropped
        assert same_as_oracle(e, ref, val, root, p) & EDGE
        p = pred.copy(); p[hub2] = 1 if root != 1 else 2
        same_as_oracle(e, ref, val, root, p)


def test_depths_past_16_bits():
    """On a path of 70000 vertices the depth map saturates at 65535 exactly
    as the reference's 16-bit map does, and the map is still valid."""
    n = 70000
    e = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int64)
    e, ref, g, val = uploaded(e, "path")
    pred = np.full(ref.n_rows, -1, np.int64)
    pred[0] = 0
    pred[1:n] = np.arange(n - 1)
    assert same_as_oracle(e, ref, val, 0, pred) == 0
    pred2 = pred.copy(); pred2[n - 1] = n - 3  # not adjacent, beyond the 16-bit range
    assert same_as_oracle(e, ref, val, 0, pred2) & PARENT
    mid = 300  # a root near one end: the far end lies past 65534 again
    p, lev, _ = O.bfs_bin(ref, mid)
    assert same_as_oracle(e, ref, val, mid, p) == 0


def test_parent_cycle_away_from_root_matches_reference():
    e = np.array([[0, 1], [1, 2], [3, 4]], np.int64)
    e, ref, g, val = uploaded(e, "cycle")
    pred = np.full(ref.n_rows, -1, np.int64)
    pred[:5] = [0, 0, 1, 4, 3]
    assert same_as_oracle(e, ref, val, 0, pred, nglobal=5) == 0


def test_device_graph_and_device_preds():
    """A device-built scale-16 graph: single-root and every 64th-lane batched
    result validated from device memory, nedge equal to the searches' own."""
    g = g5.Graph.kronecker(16)
    val = g5.Validator(g)
    roots = g.select_roots(g.nglobal)
    d = g5.DeviceArray(g.n_rows, np.int64)
    b = g5.BFS(g)
    b.run(int(roots[0]))
    b.pred_device(d)
    bad, ne, _, ms = val.validate_device(int(roots[0]), d)
    assert bad == 0 and ne == b.nedge() and ms > 0
    m = g5.MSBFS(g)
    m.run(roots)
    nedge = m.nedge()
    for lane in (0, 31, 63):
        m.pred_device(lane, d)
        bad, ne, _, _ = val.validate_device(int(roots[lane]), d)
        assert bad == 0 and ne == nedge[lane]
    p = d.to_host()
    p[int(np.nonzero(p >= 0)[0][5])] = -1
    bad, _, _, _ = val.validate(int(roots[63]), p)
    assert bad != 0


def test_rejects_a_2d_block():
    blocks = g5.Graph.kronecker_2d(10, 2, 2)
    with pytest.raises(g5.G5Error):
        g5.Validator(blocks[1])


def test_edgeless_graph():
    """No edges at all: every search reaches only its root, in every path."""
    rowptr = np.zeros(65, np.int64)
    g = g5.Graph.upload(rowptr, np.zeros(0, np.uint32), 64)
    b = g5.BFS(g)
    nl, _ = b.run(5)
    pred, lev = b.result()
    want = np.full(64, -1, np.int64)
    want[5] = 5
    assert nl == 1 and np.array_equal(pred, want) and lev[5] == 0 and (lev >= 0).sum() == 1
    assert b.nedge() == 0
    m = g5.MSBFS(g)
    m.run(np.array([0, 7, 7], np.int64))
    for lane, r in enumerate((0, 7, 7)):
        p, lv = m.result(lane)
        assert p[r] == r and (p >= 0).sum() == 1 and lv[r] == 0
    val = g5.Validator(g)
    bad, ne, depth, _ = val.validate(5, pred, nglobal=64, want_depth=True)
    assert bad == 0 and ne == 0 and depth[5] == 0 and (depth >= 0).sum() == 1
