"""ctypes bindings for the CPU oracle (oracle/liboracle.so) used by the tests.

The oracle is the checker: tests call it to produce expected values and never
route product results through it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

c_i64 = ctypes.c_int64
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p


def ptr(a):
    return None if a is None else a.ctypes.data_as(c_vp)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        src = os.path.join(ROOT, "oracle", "g5oracle.c")
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
        L = ctypes.CDLL(path)
        for name in ("orc_num_vertices", "orc_padded_rows", "orc_build_csr", "orc_nedge"):
            getattr(L, name).restype = c_i64
        for name in ("orc_s4bp128d4_max_words",):
            getattr(L, name).restype = c_sz
        _LIB = L
    return _LIB


def kronecker_edges(scale, edgefactor=16, seed1=1, seed2=2, start=0, end=None):
    L = lib()
    m = edgefactor << scale
    end = m if end is None else end
    seed = np.zeros(5, np.uint32)
    L.orc_make_mrg_seed(ctypes.c_uint64(seed1), ctypes.c_uint64(seed2), ptr(seed))
    e = np.zeros(2 * (end - start), np.int64)
    L.orc_generate_kronecker_range(ptr(seed), ctypes.c_int(scale), c_i64(start), c_i64(end), ptr(e))
    return e.reshape(-1, 2)


class CSR:
    def __init__(self, rowptr, cols, n_rows, gdeg, nglobal):
        self.rowptr, self.cols, self.n_rows, self.gdeg, self.nglobal = rowptr, cols, n_rows, gdeg, nglobal

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def build_csr(edges, n_rows=None):
    L = lib()
    e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    m = len(e) // 2
    nglobal = L.orc_num_vertices(ptr(e), c_i64(m))
    if n_rows is None:
        n_rows = L.orc_padded_rows(c_i64(nglobal))
    if n_rows < nglobal:
        raise ValueError(f"build_csr: n_rows {n_rows} < {nglobal} vertices")
    rowptr = np.zeros(n_rows + 1, np.int64)
    cols = np.zeros(max(2 * m, 1), np.uint32)
    gdeg = np.zeros(n_rows, np.int64)
    nnz = L.orc_build_csr(ptr(e), c_i64(m), c_i64(n_rows), ptr(rowptr), ptr(cols), ptr(gdeg))
    return CSR(rowptr, cols[:nnz].copy(), n_rows, gdeg, nglobal)


def bfs_bin(g, root):
    L = lib()
    pred = np.zeros(g.n_rows, np.int64)
    lev = np.zeros(g.n_rows, np.int32)
    nl = ctypes.c_int32()
    cols = g.cols if len(g.cols) else np.zeros(1, np.uint32)
    rc = L.orc_bfs_bin(ptr(g.rowptr), ptr(cols), c_i64(g.n_rows), c_i64(root), ptr(pred), ptr(lev),
                       ctypes.byref(nl))
    assert rc == 0, rc
    return pred, lev, nl.value


def bfs_simple(g, root):
    L = lib()
    pred = np.zeros(g.n_rows, np.int64)
    lev = np.zeros(g.n_rows, np.int32)
    nl = ctypes.c_int32()
    cols = g.cols if len(g.cols) else np.zeros(1, np.uint32)
    rc = L.orc_bfs_simple(ptr(g.rowptr), ptr(cols), c_i64(g.n_rows), c_i64(root), ptr(pred),
                          ptr(lev), ctypes.byref(nl))
    assert rc == 0, rc
    return pred, lev, nl.value


def validate(edges, g, nglobalverts, root, pred):
    L = lib()
    e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    depth = np.zeros(g.n_rows, np.int32)
    nedge = c_i64()
    bad = L.orc_validate(ptr(e), c_i64(len(e) // 2), c_i64(g.n_rows), c_i64(nglobalverts), c_i64(root),
                         ptr(np.ascontiguousarray(pred, dtype=np.int64)), ptr(depth), ctypes.byref(nedge))
    return bad, depth, nedge.value


def nedge(g, level):
    return lib().orc_nedge(ptr(g.gdeg), ptr(np.ascontiguousarray(level, dtype=np.int32)), c_i64(g.n_rows))


def select_roots(g, vertices, nroots=64):
    roots = np.zeros(nroots, np.int64)
    k = lib().orc_select_roots(ptr(g.rowptr), c_i64(g.n_rows), c_i64(vertices), ctypes.c_int(nroots), ptr(roots))
    return roots[:k]


def knuth_draws(vertices, count):
    out = np.zeros(count, np.int64)
    lib().orc_knuth_b_draws(c_i64(vertices), ctypes.c_int(count), ptr(out))
    return out


def aligned_zeros(n, dtype=np.uint32):
    raw = np.zeros(n + 16, dtype=dtype)
    off = (-raw.ctypes.data % 16) // raw.itemsize
    return raw[off:off + n]


def s4_encode(x):
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = aligned_zeros(L.orc_s4bp128d4_max_words(c_sz(len(x))) + 4)
    n = c_sz()
    rc = L.orc_s4bp128d4_encode(ptr(x) if len(x) else ptr(np.zeros(1, np.uint32)), c_sz(len(x)), ptr(out), ctypes.byref(n))
    assert rc == 0
    return out[:n.value].copy()


def s4_decode(words, cap):
    L = lib()
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros(cap + 128, np.uint32)
    n = c_sz()
    rc = L.orc_s4bp128d4_decode(ptr(w), c_sz(len(w)), ptr(out), ctypes.byref(n))
    return rc, out[:n.value].copy()


def cpusimd_compress(q64, threshold):
    L = lib()
    q = np.ascontiguousarray(q64, dtype=np.int64)
    out = aligned_zeros(L.orc_s4bp128d4_max_words(c_sz(len(q))) + 4)
    cs = c_sz()
    rc = L.orc_cpusimd_compress(ptr(q) if len(q) else ptr(np.zeros(1, np.int64)), c_sz(len(q)), c_sz(threshold), ptr(out), ctypes.byref(cs))
    assert rc == 0
    return out[:cs.value].copy()


def cpusimd_decompress(words, orig_n, threshold):
    L = lib()
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros(max(orig_n, len(w)) + 128, np.int64)
    n = c_sz(orig_n)
    rc = L.orc_cpusimd_decompress(ptr(w) if len(w) else ptr(np.zeros(1, np.uint32)), c_sz(len(w)), c_sz(threshold), ptr(out), ctypes.byref(n))
    assert rc == 0
    return out[:n.value].copy()


def vreduce_or(maps):
    maps = np.ascontiguousarray(maps, dtype=np.uint64)
    out = np.zeros(maps.shape[1], np.uint64)
    lib().orc_vreduce_or(ctypes.c_int(maps.shape[0]), ptr(maps), c_i64(maps.shape[1]), ptr(out))
    return out


def vreduce_min_pred(preds):
    preds = np.ascontiguousarray(preds, dtype=np.int64)
    out = np.zeros(preds.shape[1], np.int64)
    lib().orc_vreduce_min_pred(ctypes.c_int(preds.shape[0]), ptr(preds), c_i64(preds.shape[1]), ptr(out))
    return out


def golden(name):
    return os.path.join(ROOT, "tests", "golden", name)
