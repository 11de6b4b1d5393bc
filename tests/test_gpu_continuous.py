"""Continuous mode of the batched search (g5_msbfs_open/join/step/leave):
roots join between levels and lanes are reused, and every lane's result is
the oracle's single-root search of its root."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def mix64(x):
    """SplitMix64's finaliser on uint64 arrays (wrapping arithmetic)."""
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xbf58476d1ce4e5b9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94d049bb133111eb)
        x ^= x >> np.uint64(31)
    return x


def checksum(pred):
    v = np.flatnonzero(pred >= 0).astype(np.uint64)
    p1 = (pred[pred >= 0] + 1).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        return int(np.sum(mix64((v << np.uint64(32)) | p1), dtype=np.uint64))


@pytest.fixture(scope="module", params=[12, 14])
def setup(request):
    scale = request.param
    o = O.build_csr(O.kronecker_edges(scale, 16))
    g = g5.Graph.kronecker(scale)
    m = g5.MSBFS(g)
    rng = np.random.default_rng(scale)
    deg = np.diff(o.rowptr)
    cand = np.flatnonzero(deg > 0)
    roots = rng.choice(cand, 150, replace=False)
    return scale, o, g, m, roots


def _check_lane(o, m, lane, root, ne, ck):
    pred, lev = m.result(lane)
    rp, rl, _ = O.bfs_bin(o, int(root))
    n = min(len(rp), len(pred))
    assert np.array_equal(pred[:n], rp[:n]), f"root {root}"
    assert np.array_equal(lev[:n], rl[:n]), f"root {root}"
    assert ne == O.nedge(o, rl)
    assert ck == checksum(rp)


def test_staggered_joins_and_lane_reuse(setup):
    scale, o, g, m, roots = setup
    m.open()
    lane_root = {}
    pending = list(roots)
    free = list(range(64))
    done = 0
    steps = 0
    while pending or lane_root:
        # join up to 7 roots per level into free lanes
        k = min(7, len(pending), len(free))
        if k:
            lanes = [free.pop(0) for _ in range(k)]
            rs = [pending.pop(0) for _ in range(k)]
            m.join(rs, lanes)
            lane_root.update(zip(lanes, rs))
        live, _ = m.step()
        steps += 1
        finished = [ln for ln in lane_root if not (live >> ln) & 1]
        if finished:
            results = {}
            for ln in finished:
                results[ln] = m.result(ln)
            mask = 0
            for ln in finished:
                mask |= 1 << ln
            ne, ck = m.leave(mask)
            for ln in finished:
                root = lane_root.pop(ln)
                rp, rl, _ = O.bfs_bin(o, int(root))
                pred, lev = results[ln]
                n = min(len(rp), len(pred))
                assert np.array_equal(pred[:n], rp[:n]), f"root {root}"
                assert np.array_equal(lev[:n], rl[:n]), f"root {root}"
                assert int(ne[ln]) == O.nedge(o, rl)
                assert int(ck[ln]) == checksum(rp)
                free.append(ln)
                done += 1
        assert steps < 10000
    assert done == len(roots)


@pytest.mark.parametrize("lane_ratio", [0, 64, 4096])
def test_late_joiners_pushed_while_cohort_pulls(setup, lane_ratio):
    # 48 roots start together; once they run pull levels, 8 more join: a
    # split level pushes the late lanes while the cohort keeps pulling
    scale, o, g, m, roots = setup
    m.tune(lane_ratio=lane_ratio)
    m.open()
    first = [int(r) for r in roots[:48]]
    late = [int(r) for r in roots[48:56]]
    m.join(first, list(range(48)))
    lane_root = dict(zip(range(48), first))
    results = {}
    steps = 0
    while lane_root.keys() - results.keys():
        live, _ = m.step()
        steps += 1
        if steps == 2:
            m.join(late, list(range(48, 56)))
            lane_root.update(zip(range(48, 56), late))
            live |= sum(1 << ln for ln in range(48, 56))
        for ln in lane_root:
            if ln not in results and not (live >> ln) & 1:
                results[ln] = m.result(ln)
        assert steps < 1000
    split = m.split_levels()
    mask = sum(1 << ln for ln in lane_root)
    ne, ck = m.leave(mask)
    for ln, root in lane_root.items():
        rp, rl, _ = O.bfs_bin(o, root)
        pred, lev = results[ln]
        n = min(len(rp), len(pred))
        assert np.array_equal(pred[:n], rp[:n]) and np.array_equal(lev[:n], rl[:n]), f"lane {ln} root {root}"
        assert int(ck[ln]) == checksum(rp) and int(ne[ln]) == O.nedge(o, rl)
    if lane_ratio == 0:
        assert split == 0
    if lane_ratio == 64:
        assert split > 0
    m.tune(lane_ratio=64)


def test_isolated_and_duplicate_roots(setup):
    scale, o, g, m, roots = setup
    deg = np.diff(o.rowptr)
    iso = np.flatnonzero(deg == 0)
    iso = [int(iso[0])] if len(iso) else []
    rs = iso + [int(roots[0]), int(roots[0]), int(roots[1])]
    lanes = [5, 9, 10, 63][:len(rs)]
    m.open()
    m.join(rs, lanes)
    busy = sum(1 << ln for ln in lanes)
    results = {}
    for _ in range(1000):
        live, _ = m.step()
        for ln in lanes:
            if ln not in results and not (live >> ln) & 1:
                results[ln] = m.result(ln)
        if not (live & busy):
            break
    ne, ck = m.leave(busy)
    for ln, root in zip(lanes, rs):
        rp, rl, _ = O.bfs_bin(o, root)
        pred, lev = results[ln]
        n = min(len(rp), len(pred))
        assert np.array_equal(pred[:n], rp[:n]) and np.array_equal(lev[:n], rl[:n])
        assert int(ck[ln]) == checksum(rp) and int(ne[ln]) == O.nedge(o, rl)


def test_continuous_errors(setup):
    scale, o, g, m, roots = setup
    m.open()
    m.join([int(roots[0])], [3])
    with pytest.raises(g5.G5Error):
        m.join([int(roots[1])], [3])           # lane in use
    with pytest.raises(g5.G5Error):
        m.join([int(roots[1])], [64])          # lane out of range
    with pytest.raises(g5.G5Error):
        m.leave(1 << 4)                        # lane not in use
    with pytest.raises(g5.G5Error):
        m.join([g.n_rows + 5], [4])            # root out of range
    m.leave(1 << 3)


def test_batch_run_after_continuous(setup):
    """A batch run after continuous use gives the batch results again."""
    scale, o, g, m, roots = setup
    m.open()
    m.join([int(roots[2])], [0])
    m.step()
    nl, ms = m.run(roots[:8])
    pred, lev = m.result(3)
    rp, rl, _ = O.bfs_bin(o, int(roots[3]))
    n = min(len(rp), len(pred))
    assert np.array_equal(pred[:n], rp[:n]) and np.array_equal(lev[:n], rl[:n])
