"""Device Kronecker generation and CSR construction vs the oracle (which is
pinned to the reference generator's own output)."""
import numpy as np
import pytest

import gpugraph500_amd as g5
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale", [5, 10, 16])
def test_edges_match_oracle(scale):
    m = 16 << scale
    got = g5.kronecker_edges(scale, 0, m)
    assert np.array_equal(got, O.kronecker_edges(scale))


def test_edge_range_large_scale():
    # a slice of the scale-26 list: exercises 4-byte edge indices
    s, e = (1 << 30) - 5000, 1 << 30
    got = g5.kronecker_edges(26, s, e)
    assert np.array_equal(got, O.kronecker_edges(26, start=s, end=e))


@pytest.mark.parametrize("scale", [6, 12, 16])
def test_csr_matches_oracle(scale):
    g = g5.Graph.kronecker(scale)
    ref = O.build_csr(O.kronecker_edges(scale))
    rp, cols, gdeg = g.download()
    assert g.n_rows == ref.n_rows and g.nglobal == ref.nglobal
    assert g.n_gen_edges == 16 << scale
    assert np.array_equal(rp, ref.rowptr)
    assert np.array_equal(cols, ref.cols)
    assert np.array_equal(gdeg, ref.gdeg)


def test_upload_rejects_bad_csr():
    with pytest.raises(g5.G5Error):
        g5.Graph.upload(np.array([0, 1], np.int64), np.array([70], np.uint32), 64)  # col out of range
    with pytest.raises(g5.G5Error):
        g5.Graph.upload(np.zeros(11, np.int64), np.zeros(0, np.uint32), 10)  # rows not multiple of 64


@pytest.mark.parametrize("scale", [10, 16])
def test_select_roots_matches_oracle(scale):
    g = g5.Graph.kronecker(scale)
    ref = O.build_csr(O.kronecker_edges(scale))
    assert np.array_equal(g.select_roots(1 << scale, 64), O.select_roots(ref, 1 << scale))
