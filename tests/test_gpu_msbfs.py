"""Batched multi-source BFS: every lane's levels and parents bit-exact
against a separate single-root oracle search (minimum-id parent rule)."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu

_CACHE = {}


def kron(scale):
    if scale not in _CACHE:
        e = O.kronecker_edges(scale)
        ref = O.build_csr(e)
        g = g5.Graph.upload(ref.rowptr, ref.cols, ref.n_rows, gdeg=ref.gdeg)
        _CACHE[scale] = (e, ref, g)
    return _CACHE[scale]


@pytest.mark.parametrize("scale,mode", [(10, 0), (14, 0), (16, 0), (14, 1), (14, 2), (16, 2)])
def test_64_lanes_bit_exact(scale, mode):
    e, ref, g = kron(scale)
    roots = O.select_roots(ref, 1 << scale)
    m = g5.MSBFS(g)
    m.set_mode(mode)
    nl, ms = m.run(roots)
    ne = m.nedge()
    maxl = 0
    for lane, r in enumerate(roots.tolist()):
        pred, lev = m.result(lane)
        opred, olev, onl = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev), f"levels differ, lane {lane} root {r}"
        assert np.array_equal(pred, opred), f"parents differ, lane {lane} root {r}"
        assert ne[lane] == O.nedge(ref, olev)
        maxl = max(maxl, onl)
    assert nl == maxl


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_fewer_lanes_and_repeat_runs(mode):
    e, ref, g = kron(12)
    roots = O.select_roots(ref, 1 << 12)
    m = g5.MSBFS(g)
    m.set_mode(mode)
    for k in (1, 7, 33):
        sub = roots[-k:]
        m.run(sub)
        for lane, r in enumerate(sub.tolist()):
            pred, lev = m.result(lane)
            opred, olev, _ = O.bfs_bin(ref, r)
            assert np.array_equal(lev, olev) and np.array_equal(pred, opred)


def test_duplicate_roots_share_a_vertex():
    e, ref, g = kron(10)
    r = int(O.select_roots(ref, 1 << 10)[0])
    m = g5.MSBFS(g)
    m.run([r, r, r])
    opred, olev, _ = O.bfs_bin(ref, r)
    for lane in range(3):
        pred, lev = m.result(lane)
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred)


def test_mixed_roots_keep_the_callers_lane_order():
    # lanes are ranked by the roots' reach inside a batch run; results come
    # back in the caller's order, also with an edgeless root, duplicates and
    # ties in reach
    e, ref, g = kron(12)
    deg = np.diff(ref.rowptr)
    iso = np.flatnonzero(deg == 0)
    sel = O.select_roots(ref, 1 << 12)
    rs = [int(sel[5]), int(sel[0]), int(sel[5])] + ([int(iso[0])] if len(iso) else []) + [int(x) for x in sel[10:40]]
    m = g5.MSBFS(g)
    m.run(rs)
    ne = m.nedge()
    for lane, r in enumerate(rs):
        pred, lev = m.result(lane)
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred), (lane, r)
        assert ne[lane] == O.nedge(ref, olev)
        d = g5.DeviceArray(g.n_cols, np.int64)
        m.pred_device(lane, d)
        assert np.array_equal(d.to_host(), opred)


def test_reach_is_the_neighbours_edges():
    e, ref, g = kron(12)
    deg = np.diff(ref.rowptr)
    roots = np.concatenate([O.select_roots(ref, 1 << 12)[:40], np.flatnonzero(deg == 0)[:2]])
    m = g5.MSBFS(g)
    want = [int(deg[ref.cols[ref.rowptr[r]:ref.rowptr[r + 1]]].sum()) for r in roots.tolist()]
    assert m.reach(roots).tolist() == want
    with pytest.raises(g5.G5Error):
        m.reach(np.arange(65))


def test_profile_path_equals_graph_path():
    e, ref, g = kron(14)
    roots = O.select_roots(ref, 1 << 14)
    m = g5.MSBFS(g)
    m.profile(True)
    m.run(roots)
    st = m.stats()
    assert sum(v["launches"] for v in st.values()) > 0
    pa = [m.result(l) for l in (0, 31, 63)]
    m.profile(False)
    m.run(roots)
    pb = [m.result(l) for l in (0, 31, 63)]
    for (p1, l1), (p2, l2) in zip(pa, pb):
        assert np.array_equal(p1, p2) and np.array_equal(l1, l2)


def test_tiny_graphs_all_roots_in_one_batch():
    rng = np.random.default_rng(77)
    for trial in range(20):
        n = int(rng.integers(2, 60))
        m_ = int(rng.integers(1, 5 * n))
        edges = rng.integers(0, n, size=(m_, 2))
        # add one hub so the wave and block kernels see work
        hub = np.stack([np.zeros(3000, np.int64), rng.integers(0, n, 3000)], 1) if trial % 5 == 0 else np.zeros((0, 2), np.int64)
        ref = O.build_csr(np.concatenate([edges, hub]), 64)
        g = g5.Graph.upload(ref.rowptr, ref.cols, 64, gdeg=ref.gdeg)
        roots = [v for v in range(n) if ref.rowptr[v + 1] > ref.rowptr[v]][:64]
        if not roots:
            continue
        ms = g5.MSBFS(g)
        for mode in (0, 1, 2):
            ms.set_mode(mode)
            ms.run(roots)
            for lane, r in enumerate(roots):
                pred, lev = ms.result(lane)
                opred, olev, _ = O.bfs_bin(ref, r)
                assert np.array_equal(lev, olev) and np.array_equal(pred, opred), (trial, mode, r)
        ms.free()
        g.free()


def test_bad_roots_rejected():
    e, ref, g = kron(10)
    m = g5.MSBFS(g)
    with pytest.raises(g5.G5Error):
        m.run(np.arange(65))
    with pytest.raises(g5.G5Error):
        m.run([ref.n_rows])


def _check_all_lanes(m, ref, roots):
    ne = m.nedge()
    for lane, r in enumerate(roots.tolist()):
        pred, lev = m.result(lane)
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev), f"levels differ, lane {lane} root {r}"
        assert np.array_equal(pred, opred), f"parents differ, lane {lane} root {r}"
        assert ne[lane] == O.nedge(ref, olev)


@pytest.mark.parametrize("shape", [
    dict(thread_max=0, wave_max=0),              # every vertex in the workgroup kernel
    dict(thread_max=1 << 20, wave_max=1 << 20),  # every vertex in the thread kernel
    dict(thread_max=0, wave_max=1 << 20),        # every vertex in the wavefront kernel
    dict(thread_max=4, wave_max=64, push_ratio=1, active_ratio=1),
    dict(thread_max=3, wave_max=100),            # bounds inside degree classes: mixed bins
    dict(push_ratio=1000, active_ratio=1000),
    dict(lane_ratio=0),                          # no split levels
    dict(lane_ratio=4, push_ratio=2),            # split levels push most lanes
])
def test_shape_parameters_do_not_change_results(shape):
    e, ref, g = kron(14)
    roots = O.select_roots(ref, 1 << 14)
    m = g5.MSBFS(g)
    m.tune(**shape)
    got = m.get_tune()
    for k, v in shape.items():
        assert got[k] == v
    m.run(roots)
    _check_all_lanes(m, ref, roots)


@pytest.mark.parametrize("scale,shape", [
    (12, {}), (14, {}),
    # every vertex in the thread kernel: a split level pushes into no bin,
    # the thread kernel pulls every lane
    (14, dict(thread_max=1 << 20, wave_max=1 << 20, split_max=1 << 20)),
    # no thread bin: the split level's push reaches every target
    (14, dict(thread_max=0, wave_max=64)),
    # split levels with the pull levels' thread bin, with a quarter and an
    # eighth of it (the level after a split then trusts the level flags only
    # from the split level's thread bin on), and with none
    (14, dict(split_max=32)),
    (14, dict(thread_max=64, split_max=16)),
    (14, dict(thread_max=256, split_max=32)),
    (14, dict(split_max=0)),
    # a small push budget: the split takes the lanes with the fewest edges
    # that fit in it, not all candidates or none
    (14, dict(push_ratio=64)),
])
def test_split_levels_match_oracle(scale, shape):
    # a pull level after a push level pushes the lanes with few frontier
    # edges (into the block and wave bins; the thread kernel pulls them);
    # every lane ratio gives the oracle's levels and parents, and some ratio
    # splits at least one level.  Few pushed lanes on few block and wave
    # vertices take the candidate pass, many (or no thread bin) the two
    # passes of a push level.
    e, ref, g = kron(scale)
    roots = O.select_roots(ref, 1 << scale)
    m = g5.MSBFS(g)
    if shape:
        m.tune(**shape)
    m.profile(True)
    kinds = set()
    for ratio in (16, 64, 256, 1024):
        m.tune(lane_ratio=ratio)
        m.run(roots)
        kinds |= {int(r[9]) for r in m.levels()}
        _check_all_lanes(m, ref, roots)
    assert 2 in kinds, kinds
    m.profile(False)
    m.tune(lane_ratio=64)
    m.run(roots)           # the graph-replayed pull levels with a lane mask
    _check_all_lanes(m, ref, roots)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_back_to_back_runs(mode):
    # a run after a run whose last level was a push level clears the frontier
    # buffers entry by entry; after a pull level, whole
    e, ref, g = kron(13)
    roots = O.select_roots(ref, 1 << 13)
    m = g5.MSBFS(g)
    m.set_mode(mode)
    for rs in (roots, roots[::-1][:40], roots[5:9], roots):
        m.run(rs)
        _check_all_lanes(m, ref, rs)
    # continuous steps in between, then a batch run again
    m.open()
    m.join(roots[:3].tolist(), [0, 1, 2])
    m.step()
    m.run(roots[10:30])
    _check_all_lanes(m, ref, roots[10:30])


def test_autotune_keeps_results():
    e, ref, g = kron(13)
    roots = O.select_roots(ref, 1 << 13)
    m = g5.MSBFS(g)
    best, trials, params = m.autotune(roots)
    assert best > 0 and trials > 10
    assert params == m.get_tune()
    m.run(roots)
    _check_all_lanes(m, ref, roots)


def test_tune_rejects_bad_values():
    e, ref, g = kron(10)
    m = g5.MSBFS(g)
    for bad in (dict(thread_max=-1), dict(thread_max=100, wave_max=50), dict(push_ratio=0), dict(active_ratio=0),
                dict(lane_ratio=-1), dict(split_max=-1)):
        m2 = g5.MSBFS(g)
        with pytest.raises(g5.G5Error):
            m2.tune(**bad)
    with pytest.raises(g5.G5Error):
        g5.check(g5.lib().g5_msbfs_tune(m.h, 9, 1))


def _download_maps(m, n):
    dp, dl, lane_of, final = m.maps()
    par = np.empty((n, 64), np.int32)
    lev = np.empty((n, 64), np.uint8)
    g5.check(g5.lib().g5_memcpy_d2h(par.ctypes.data_as(g5.c_vp), g5.c_vp(dp), par.nbytes))
    g5.check(g5.lib().g5_memcpy_d2h(lev.ctypes.data_as(g5.c_vp), g5.c_vp(dl), lev.nbytes))
    return par, lev, lane_of, final


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_final_maps_as_they_lie_in_hbm(mode):
    # the run's own output, with no extraction: row v (original id), slot
    # lane_of[i] = root i's parent and level, -1 / 255 where unreached
    e, ref, g = kron(12)
    roots = O.select_roots(ref, 1 << 12)
    m = g5.MSBFS(g)
    m.set_mode(mode)
    for rs in (roots, roots[7:20]):
        m.run(rs)
        par, lev, lane_of, final = _download_maps(m, g.n_rows)
        assert final
        assert sorted(lane_of.tolist()) == list(range(len(rs)))
        for i, r in enumerate(rs.tolist()):
            opred, olev, _ = O.bfs_bin(ref, r)
            assert np.array_equal(par[:, lane_of[i]].astype(np.int64), opred), (mode, i)
            want = np.where(olev < 0, 255, olev).astype(np.uint8)
            assert np.array_equal(lev[:, lane_of[i]], want), (mode, i)


def test_record_placement_keeps_results():
    # fresh allocations of the parent records (copies) keep every map right:
    # the kept placement holds the placement roots' maps, later runs (an
    # edgeless root among the earlier ones) the oracle's
    e, ref, g = kron(12)
    deg = np.diff(ref.rowptr)
    iso = np.flatnonzero(deg == 0)
    sel = O.select_roots(ref, 1 << 12)
    m = g5.MSBFS(g)
    m.run([int(iso[0])] + sel[:20].tolist())
    first, best, moves = m.place(sel[20:60], trials=4)
    assert 0 < best <= first and 0 <= moves <= 4
    par, lev, lane_of, final = _download_maps(m, g.n_rows)
    assert final
    for i, r in enumerate(sel[20:60].tolist()):
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(par[:, lane_of[i]].astype(np.int64), opred), i
        assert np.array_equal(lev[:, lane_of[i]], np.where(olev < 0, 255, olev).astype(np.uint8)), i
    m.run(sel)
    _check_all_lanes(m, ref, sel)
    par, lev, lane_of, _ = _download_maps(m, g.n_rows)
    assert (par[iso[0], :] == -1).all() and (lev[iso[0], :] == 255).all()
    with pytest.raises(g5.G5Error):
        m.place(sel, trials=17)
    m.open()
    m.join([int(sel[0])], [0])
    with pytest.raises(g5.G5Error):
        m.place(sel, trials=1)


def test_edgeless_root_leaves_no_trace():
    # an edgeless root's own slot holds the root; the next runs, with other
    # roots in that lane, read the vertex as unreached again (also after
    # continuous-mode joins of edgeless roots)
    e, ref, g = kron(12)
    deg = np.diff(ref.rowptr)
    iso = np.flatnonzero(deg == 0)
    assert len(iso) >= 3
    sel = O.select_roots(ref, 1 << 12)
    m = g5.MSBFS(g)
    m.run([int(iso[0])] + sel[:10].tolist())
    par, lev, lane_of, _ = _download_maps(m, g.n_rows)
    assert par[iso[0], lane_of[0]] == iso[0] and lev[iso[0], lane_of[0]] == 0
    m.open()
    m.join([int(iso[1]), int(sel[3])], [0, 5])
    m.step()
    m.run(sel[20:60])
    par, lev, lane_of, _ = _download_maps(m, g.n_rows)
    used = lane_of.tolist()
    for v in iso[:3]:
        assert (par[v, used] == -1).all() and (lev[v, used] == 255).all()
    _check_all_lanes(m, ref, sel[20:60])


def test_deep_search_levels_from_the_parent_tree():
    # a path of 600 vertices: 599 levels do not fit the level bytes, so the
    # run reports them as not final and the per-root result derives levels
    # from the parent tree; a star beside it (centre 600) is shallow
    n = 600
    path = np.stack([np.arange(n - 1), np.arange(1, n)], 1)
    star = np.stack([np.full(39, 600), np.arange(601, 640)], 1)
    ref = O.build_csr(np.concatenate([path, star]), 640)
    g = g5.Graph.upload(ref.rowptr, ref.cols, 640, gdeg=ref.gdeg)
    m = g5.MSBFS(g)
    roots = [0, 299, 599]
    nl, _ = m.run(roots)
    assert nl == 600
    _, _, _, final = m.maps()
    assert not final
    for lane, r in enumerate(roots):
        pred, lev = m.result(lane)
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(pred, opred) and np.array_equal(lev, olev), r
        d = g5.DeviceArray(640, np.int32)
        with pytest.raises(g5.G5Error):
            m.level_device(lane, d)
    # a shallow run afterwards has final level bytes again
    m.run([600, 605])
    assert m.maps()[3]
    for lane, r in enumerate([600, 605]):
        pred, lev = m.result(lane)
        opred, olev, _ = O.bfs_bin(ref, r)
        assert np.array_equal(lev, olev) and np.array_equal(pred, opred)
        d = g5.DeviceArray(640, np.int32)
        m.level_device(lane, d)
        assert np.array_equal(d.to_host(), olev)
