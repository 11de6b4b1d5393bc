"""Root selection (cpu_2d/main.cpp:351-411) pinned by libstdc++'s own
std::knuth_b/uniform_int_distribution draws, and the vreduce combines pinned
by numpy closed forms."""
import json

import numpy as np
import pytest

import oracle_lib as O

with open(O.golden("roots_draws.json")) as f:
    DRAWS = json.load(f)


@pytest.mark.parametrize("scale", sorted(int(s) for s in DRAWS))
def test_knuth_b_uniform_draws(scale):
    want = np.array(DRAWS[str(scale)], np.int64)
    assert np.array_equal(O.knuth_draws(1 << scale, len(want)), want)


def test_select_roots_rule():
    e = O.kronecker_edges(10)
    g = O.build_csr(e)
    roots = O.select_roots(g, 1 << 10)
    draws = np.array(DRAWS["10"], np.int64)
    # roots are the first distinct draws with nonzero CSR degree, in order
    want, seen = [], set()
    for d in draws.tolist():
        if len(want) == 64:
            break
        if d in seen:
            continue
        seen.add(d)
        if d < g.n_rows and g.rowptr[d + 1] > g.rowptr[d]:
            want.append(d)
    assert roots.tolist() == want


def test_vreduce_closed_forms():
    rng = np.random.default_rng(5)
    maps = rng.integers(0, 1 << 63, size=(4, 1000), dtype=np.uint64) & rng.integers(0, 1 << 63, size=(4, 1000), dtype=np.uint64)
    assert np.array_equal(O.vreduce_or(maps), np.bitwise_or.reduce(maps, axis=0))
    preds = rng.integers(-1, 500, size=(4, 1000)).astype(np.int64)
    masked = np.where(preds < 0, np.iinfo(np.int64).max, preds).min(axis=0)
    want = np.where(masked == np.iinfo(np.int64).max, -1, masked)
    assert np.array_equal(O.vreduce_min_pred(preds), want)


def test_tuning_roots_follow_the_timed_roots():
    # the benchmark tunes on the next 64 draws: asking for 128 keys gives the
    # 64 timed keys first, unchanged, then 64 disjoint ones
    g = O.build_csr(O.kronecker_edges(14))
    timed = O.select_roots(g, 1 << 14, 64)
    both = O.select_roots(g, 1 << 14, 128)
    assert np.array_equal(both[:64], timed)
    assert len(both) == 128 and not set(both[64:].tolist()) & set(timed.tolist())
