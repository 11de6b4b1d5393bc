"""gpugraph500_amd -- MI355X (gfx950) hot path of gpugraph500's 2D BFS.

Python bindings (ctypes) over the C ABI declared in include/g5hip.h and
implemented by libg5hip.so (HIP, built in-tree by gpugraph500_amd.build).
There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg5hip.so")

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_u64 = ctypes.c_uint64

_lib = None


class G5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"g5 error {code}: {msg}")
        self.code = code


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def lib():
    """Load libg5hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise G5Error(-100, f"{LIB_PATH} is missing; run python -m gpugraph500_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    I = ctypes.c_int
    P = c_vp
    _sig(L, "g5_last_error", ctypes.c_char_p, [])
    _sig(L, "g5_init", I, [I])
    _sig(L, "g5_device_sync", I, [])
    _sig(L, "g5_dmalloc", I, [ctypes.POINTER(c_vp), c_sz])
    _sig(L, "g5_dfree", I, [P])
    _sig(L, "g5_memcpy_h2d", I, [P, P, c_sz])
    _sig(L, "g5_memcpy_d2h", I, [P, P, c_sz])
    _sig(L, "g5_graph_upload", I, [P, P, c_i64, c_i64, P, c_i64, c_i64, ctypes.POINTER(c_vp)])
    _sig(L, "g5_graph_kronecker", I, [I, I, c_u64, c_u64, ctypes.POINTER(c_vp)])
    _sig(L, "g5_graph_kronecker_2d", I, [I, I, c_u64, c_u64, I, I, P])
    _sig(L, "g5_block_bounds", I, [c_i64, I, I, I, I, P, P, P, P])
    _sig(L, "g5_graph_bounds", I, [P, P, P])
    _sig(L, "g5_graph_gdeg", I, [P, P])
    _sig(L, "g5_graph_degree_hist", I, [P, P, I])
    _sig(L, "g5_memcpy_d2d", I, [P, P, c_sz])
    _sig(L, "g5_fill_u64", I, [P, c_i64, c_u64])
    _sig(L, "g5_bitmap_to_queue", I, [P, c_i64, P, ctypes.POINTER(c_i64)])
    _sig(L, "g5_queue_to_bitmap", I, [P, c_i64, P, c_i64])
    _sig(L, "g5_bfs_out_bitmap", I, [P, ctypes.POINTER(c_vp), ctypes.POINTER(c_i64)])
    _sig(L, "g5_bfs_set_out_bitmap", I, [P, P])
    _sig(L, "g5_bfs_set_in_bitmap", I, [P, P])
    _sig(L, "g5_bfs_pred_device", I, [P, P])
    _sig(L, "g5_bfs_expand_bottom_up", I, [P, P, ctypes.POINTER(I)])
    _sig(L, "g5_bitmap_count", I, [P, c_i64, ctypes.POINTER(c_i64)])
    _sig(L, "g5_graph_info", I, [P, P, P, P, P, P])
    _sig(L, "g5_graph_download", I, [P, P, P, P])
    _sig(L, "g5_graph_free", I, [P])
    _sig(L, "g5_kronecker_edges", I, [I, c_u64, c_u64, c_i64, c_i64, P])
    _sig(L, "g5_select_roots", I, [P, c_i64, I, P, ctypes.POINTER(I)])
    _sig(L, "g5_bfs_profile", I, [P, I])
    _sig(L, "g5_bfs_stats", I, [P, P, P, P, P])
    _sig(L, "g5_bfs_kernel_stats", I, [P, I, P, P, P, P])
    _sig(L, "g5_bfs_create", I, [P, ctypes.POINTER(c_vp)])
    _sig(L, "g5_bfs_free", I, [P])
    _sig(L, "g5_bfs_start", I, [P, c_i64])
    _sig(L, "g5_bfs_expand", I, [P, ctypes.POINTER(I)])
    _sig(L, "g5_bfs_outqueue", I, [P, P, ctypes.POINTER(c_i64)])
    _sig(L, "g5_bfs_set_outqueue", I, [P, P, c_i64])
    _sig(L, "g5_bfs_set_inqueue", I, [P, P, c_i64])
    _sig(L, "g5_bfs_result", I, [P, P, P])
    _sig(L, "g5_bfs_run", I, [P, c_i64, I, ctypes.POINTER(c_i32), ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_bfs_tune", I, [P, c_i64, c_i64])
    _sig(L, "g5_bfs_nedge", I, [P, ctypes.POINTER(c_i64)])
    _sig(L, "g5_bfs_summary", I, [P, ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_uint64)])
    _sig(L, "g5_msbfs_create", I, [P, ctypes.POINTER(c_vp)])
    _sig(L, "g5_msbfs_free", I, [P])
    _sig(L, "g5_msbfs_run", I, [P, P, I, ctypes.POINTER(c_i32), ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_msbfs_result", I, [P, I, P, P])
    _sig(L, "g5_msbfs_pred_device", I, [P, I, P])
    _sig(L, "g5_msbfs_level_device", I, [P, I, P])
    _sig(L, "g5_msbfs_maps", I, [P, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), P, ctypes.POINTER(I)])
    _sig(L, "g5_msbfs_times", I, [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_msbfs_footprint", I, [P, P, P, P])
    _sig(L, "g5_mem_info", I, [P, P])
    _sig(L, "g5_bfs_level_device", I, [P, P])
    _sig(L, "g5_diff_u32", I, [P, P, c_i64, ctypes.POINTER(c_i64)])
    _sig(L, "g5_exchange_profile", I, [I])
    _sig(L, "g5_exchange_stats", I, [I, P, P, P])
    _sig(L, "g5_validator_create", I, [P, ctypes.POINTER(c_vp)])
    _sig(L, "g5_validator_free", I, [P])
    _sig(L, "g5_validate", I, [P, c_i64, c_i64, P, ctypes.POINTER(I), ctypes.POINTER(c_i64), P,
                               ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_validate_device", I, [P, c_i64, c_i64, P, ctypes.POINTER(I), ctypes.POINTER(c_i64), P,
                                      ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_msbfs_nedge", I, [P, P])
    _sig(L, "g5_msbfs_profile", I, [P, I])
    _sig(L, "g5_msbfs_stats", I, [P, I, P, P, P])
    _sig(L, "g5_msbfs_info", I, [P, P, I])
    _sig(L, "g5_msbfs_levels", I, [P, P, I, ctypes.POINTER(I)])
    _sig(L, "g5_msbfs_set_mode", I, [P, I])
    _sig(L, "g5_msbfs_open", I, [P])
    _sig(L, "g5_msbfs_tune", I, [P, I, c_i64])
    _sig(L, "g5_msbfs_get_tune", I, [P, I, ctypes.POINTER(c_i64)])
    _sig(L, "g5_msbfs_autotune", I, [P, P, I, I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I)])
    _sig(L, "g5_msbfs_place", I, [P, P, I, I, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(I)])
    _sig(L, "g5_msbfs_join", I, [P, P, P, I])
    _sig(L, "g5_msbfs_step", I, [P, ctypes.POINTER(c_u64), ctypes.POINTER(ctypes.c_float)])
    _sig(L, "g5_msbfs_leave", I, [P, c_u64, P, P])
    _sig(L, "g5_s4_max_words", c_sz, [c_sz])
    _sig(L, "g5_s4_encode", I, [P, c_sz, P, c_sz, ctypes.POINTER(c_sz)])
    _sig(L, "g5_s4_decode", I, [P, c_sz, P, c_sz, ctypes.POINTER(c_sz)])
    _sig(L, "g5_s4_roundtrip", I, [P, c_sz, P, c_sz, P, ctypes.POINTER(c_sz)])
    _sig(L, "g5_s4_roundtrip_noalloc", I, [P, c_sz, P, c_sz, P, ctypes.POINTER(c_sz)])
    _sig(L, "g5_vreduce_or", I, [P, I, c_i64, P])
    _sig(L, "g5_vreduce_min_pred", I, [P, I, c_i64, P])
    _sig(L, "g5_vreduce_or_list", I, [P, I, c_i64, P])
    _sig(L, "g5_msbfs_reach", I, [P, P, I, P])
    _sig(L, "g5_vreduce_min_pred_list", I, [P, I, c_i64, P])
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().g5_last_error()
        raise G5Error(rc, msg.decode() if msg else "")
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(c_vp)


def init(device=0):
    check(lib().g5_init(device))


EXCHANGE_UNITS = ("or", "min_pred", "encode", "decode")


def exchange_profile(enable=True):
    """Device-time accounting of the vreduce and codec entry points (resets
    the totals)."""
    check(lib().g5_exchange_profile(1 if enable else 0))


def exchange_stats():
    """{unit: {"ms", "bytes", "calls"}} since exchange_profile(True)."""
    out = {}
    for u, name in enumerate(EXCHANGE_UNITS):
        ms, by, calls = ctypes.c_double(), ctypes.c_double(), c_i64()
        check(lib().g5_exchange_stats(u, ctypes.byref(ms), ctypes.byref(by), ctypes.byref(calls)))
        out[name] = {"ms": ms.value, "bytes": by.value, "calls": calls.value}
    return out


def diff_u32(a, b, n_words):
    """Positions where two device buffers (DeviceArray or raw device
    addresses) differ, compared as n_words uint32 words on the device."""
    pa = a.ptr if hasattr(a, "ptr") else c_vp(a)
    pb = b.ptr if hasattr(b, "ptr") else c_vp(b)
    out = c_i64()
    check(lib().g5_diff_u32(pa, pb, int(n_words), ctypes.byref(out)))
    return out.value


class DeviceArray:
    """A device (HBM) buffer owned by the library's allocator."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.ptr = c_vp()
        check(lib().g5_dmalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().g5_memcpy_h2d(d.ptr, _p(a), a.nbytes))
        return d

    def to_host(self, n=None):
        n = self.n if n is None else int(n)
        out = np.empty(n, self.dtype)
        if n:
            check(lib().g5_memcpy_d2h(_p(out), self.ptr, n * self.dtype.itemsize))
        return out

    def free(self):
        if self.ptr:
            lib().g5_dfree(self.ptr)
            self.ptr = c_vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """One rank's DistMatrix2d block resident in HBM."""

    def __init__(self, handle):
        self.h = handle
        info = [c_i64() for _ in range(5)]
        check(lib().g5_graph_info(self.h, *[ctypes.byref(x) for x in info]))
        self.n_rows, self.n_cols, self.nnz, self.nglobal, self.n_gen_edges = [x.value for x in info]
        rs, cs = c_i64(), c_i64()
        check(lib().g5_graph_bounds(self.h, ctypes.byref(rs), ctypes.byref(cs)))
        self.row_start, self.col_start = rs.value, cs.value

    @classmethod
    def kronecker(cls, scale, edgefactor=16, seed1=1, seed2=2):
        h = c_vp()
        check(lib().g5_graph_kronecker(scale, edgefactor, seed1, seed2, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def kronecker_2d(cls, scale, R, C, edgefactor=16, seed1=1, seed2=2):
        """The R x C blocks of the graph, as a list indexed [r * C + c]."""
        hs = (c_vp * (R * C))()
        check(lib().g5_graph_kronecker_2d(scale, edgefactor, seed1, seed2, R, C, hs))
        return [cls(c_vp(h)) for h in hs]

    @classmethod
    def upload(cls, rowptr, cols, n_rows, n_cols=None, gdeg=None, row_start=0, col_start=0):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        gd = None if gdeg is None else np.ascontiguousarray(gdeg, dtype=np.int64)
        h = c_vp()
        check(lib().g5_graph_upload(_p(rowptr), _p(cols) if cols.size else None, n_rows,
                                    n_rows if n_cols is None else n_cols, _p(gd), row_start,
                                    col_start, ctypes.byref(h)))
        return cls(h)

    def select_roots(self, vertices, nroots=64):
        roots = np.empty(max(nroots, 1), np.int64)
        found = ctypes.c_int()
        check(lib().g5_select_roots(self.h, vertices, nroots, _p(roots), ctypes.byref(found)))
        return roots[:found.value].copy()

    def degree_hist(self, nbins=40):
        """Rows per degree class: [0] no neighbours, [k] 2^(k-1) <= degree < 2^k."""
        out = np.zeros(nbins, np.int64)
        check(lib().g5_graph_degree_hist(self.h, _p(out), nbins))
        return out

    def download(self):
        rp = np.empty(self.n_rows + 1, np.int64)
        cols = np.empty(max(self.nnz, 1), np.uint32)
        gdeg = np.empty(max(self.n_rows, 1), np.int64)
        check(lib().g5_graph_download(self.h, _p(rp), _p(cols), _p(gdeg)))
        return rp, cols[:self.nnz], gdeg[:self.n_rows]

    def free(self):
        if self.h:
            lib().g5_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def block_bounds(nglobal, R, C, r, c):
    """(row_start, n_rows, col_start, n_cols) of block (r, c) (host only)."""
    out = [c_i64() for _ in range(4)]
    check(lib().g5_block_bounds(nglobal, R, C, r, c, *[ctypes.byref(x) for x in out]))
    return tuple(x.value for x in out)


def kronecker_edges(scale, start, end, seed1=1, seed2=2):
    out = np.empty(2 * (end - start), np.int64)
    check(lib().g5_kronecker_edges(scale, seed1, seed2, start, end, _p(out)))
    return out.reshape(-1, 2)


class BFS:
    """Single-root local expansion (the CUDA_BFS plug-in's hooks)."""

    def __init__(self, graph):
        self.graph = graph
        self.h = c_vp()
        check(lib().g5_bfs_create(graph.h, ctypes.byref(self.h)))

    def start(self, root):
        check(lib().g5_bfs_start(self.h, root))

    def expand(self):
        a = ctypes.c_int()
        check(lib().g5_bfs_expand(self.h, ctypes.byref(a)))
        return bool(a.value)

    def expand_bottom_up(self, transposed):
        """A bottom-up level over the transposed block (a Graph); the row
        frontier must come from set_in_bitmap."""
        a = ctypes.c_int()
        check(lib().g5_bfs_expand_bottom_up(self.h, transposed.h, ctypes.byref(a)))
        return bool(a.value)

    def outqueue(self):
        q = np.empty(max(self.graph.n_cols, 1), np.uint32)
        n = c_i64(q.size)
        check(lib().g5_bfs_outqueue(self.h, _p(q), ctypes.byref(n)))
        return q[:n.value].copy()

    def set_outqueue(self, q=None):
        if q is None:
            check(lib().g5_bfs_set_outqueue(self.h, None, 0))
        else:
            q = np.ascontiguousarray(q, dtype=np.uint32)
            check(lib().g5_bfs_set_outqueue(self.h, _p(q) if q.size else None, q.size))

    def set_inqueue(self, q=None):
        if q is None:
            check(lib().g5_bfs_set_inqueue(self.h, None, 0))
        else:
            q = np.ascontiguousarray(q, dtype=np.uint32)
            check(lib().g5_bfs_set_inqueue(self.h, _p(q) if q.size else np.zeros(1, np.uint32).ctypes.data_as(c_vp), q.size))

    def result(self):
        pred = np.empty(self.graph.n_cols, np.int64)
        lev = np.empty(self.graph.n_cols, np.int32)
        check(lib().g5_bfs_result(self.h, _p(pred), _p(lev)))
        return pred, lev

    def pred_device(self, out):
        """The predecessor map (global ids) into the DeviceArray `out` (int64[n_cols])."""
        check(lib().g5_bfs_pred_device(self.h, out.ptr))

    def level_device(self, out):
        """The level map into the DeviceArray `out` (int32[n_cols])."""
        check(lib().g5_bfs_level_device(self.h, out.ptr))

    RUN_CODEC, RUN_TOPDOWN = 1, 2
    KERNELS = ("k_expand", "k_bottom_up")

    def run(self, root, codec=False, topdown=False):
        nl = c_i32()
        ms = ctypes.c_float()
        flags = (self.RUN_CODEC if codec else 0) | (self.RUN_TOPDOWN if topdown else 0)
        check(lib().g5_bfs_run(self.h, root, flags, ctypes.byref(nl), ctypes.byref(ms)))
        return nl.value, ms.value

    def tune(self, alpha=20, beta=24):
        """Direction switch of run(): bottom-up once the frontier's edges exceed
        the edges not yet checked top-down / alpha, top-down again below
        n / beta frontier vertices."""
        check(lib().g5_bfs_tune(self.h, alpha, beta))

    def profile(self, enable=True):
        check(lib().g5_bfs_profile(self.h, 1 if enable else 0))

    def stats(self):
        ms, bytes_ = ctypes.c_double(), ctypes.c_double()
        launches, edges = c_i64(), c_i64()
        check(lib().g5_bfs_stats(self.h, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(bytes_),
                                 ctypes.byref(edges)))
        return {"ms": ms.value, "launches": launches.value, "bytes": bytes_.value, "edges": edges.value}

    def kernel_stats(self):
        """{kernel name: stats} for k_expand and k_bottom_up."""
        res = {}
        for k, name in enumerate(self.KERNELS):
            ms, bytes_ = ctypes.c_double(), ctypes.c_double()
            launches, edges = c_i64(), c_i64()
            check(lib().g5_bfs_kernel_stats(self.h, k, ctypes.byref(ms), ctypes.byref(launches),
                                            ctypes.byref(bytes_), ctypes.byref(edges)))
            res[name] = {"ms": ms.value, "launches": launches.value, "bytes": bytes_.value, "edges": edges.value}
        return res

    def nedge(self):
        n = c_i64()
        check(lib().g5_bfs_nedge(self.h, ctypes.byref(n)))
        return n.value

    def summary(self):
        """(nedge, checksum of the predecessor map) of the last run, on the device."""
        n, ck = c_i64(), ctypes.c_uint64()
        check(lib().g5_bfs_summary(self.h, ctypes.byref(n), ctypes.byref(ck)))
        return n.value, ck.value

    def free(self):
        if self.h:
            lib().g5_bfs_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MSBFS:
    """Batched multi-source BFS over up to 64 roots (uint64 lane masks)."""

    KERNELS = ("k_pull_block", "k_pull_wave", "k_pull_thread", "push")

    def __init__(self, graph):
        self.graph = graph
        self.h = c_vp()
        check(lib().g5_msbfs_create(graph.h, ctypes.byref(self.h)))
        self.nroots = 0

    def run(self, roots):
        r = np.ascontiguousarray(roots, dtype=np.int64)
        nl = c_i32()
        ms = ctypes.c_float()
        check(lib().g5_msbfs_run(self.h, _p(r), r.size, ctypes.byref(nl), ctypes.byref(ms)))
        self.nroots = r.size
        return nl.value, ms.value

    def result(self, lane):
        pred = np.empty(self.graph.n_cols, np.int64)
        lev = np.empty(self.graph.n_cols, np.int32)
        check(lib().g5_msbfs_result(self.h, lane, _p(pred), _p(lev)))
        return pred, lev

    def pred_device(self, lane, out):
        """Lane `lane`'s predecessor map into the DeviceArray `out` (int64[n])."""
        check(lib().g5_msbfs_pred_device(self.h, lane, out.ptr))

    def level_device(self, lane, out):
        """Lane `lane`'s level map into the DeviceArray `out` (int32[n])."""
        check(lib().g5_msbfs_level_device(self.h, lane, out.ptr))

    def maps(self):
        """The last batch run's final maps as they lie in HBM: (device pointer
        of parent int32[n][64], device pointer of level uint8[n][64], lane of
        each root, levels_final)."""
        dp, dl, fin = c_vp(), c_vp(), ctypes.c_int()
        lane_of = np.zeros(max(self.nroots, 1), np.int32)
        check(lib().g5_msbfs_maps(self.h, ctypes.byref(dp), ctypes.byref(dl), _p(lane_of), ctypes.byref(fin)))
        return dp.value, dl.value, lane_of[:self.nroots], bool(fin.value)

    @staticmethod
    def footprint(graph):
        """(graph bytes, fixed bytes, bytes per lane) of a batched search over graph."""
        a, b, c = c_i64(), c_i64(), c_i64()
        check(lib().g5_msbfs_footprint(graph.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def times(self):
        """(search ms, finalising pass ms) of the last batch run."""
        a, b = ctypes.c_float(), ctypes.c_float()
        check(lib().g5_msbfs_times(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # ---- shape parameters ----
    TUNE_KEYS = {"thread_max": 0, "wave_max": 1, "push_ratio": 2, "active_ratio": 3, "lane_ratio": 4,
                 "split_max": 5}

    def tune(self, **params):
        """Set shape parameters; the bin bounds are applied in the order that
        keeps thread_max <= wave_max at every step."""
        keys = list(params)
        if "thread_max" in params and "wave_max" in params:
            cur_wave = self.get_tune()["wave_max"]
            first = "wave_max" if params["thread_max"] > cur_wave else "thread_max"
            keys = [first] + [k for k in keys if k != first]
        for k in keys:
            check(lib().g5_msbfs_tune(self.h, self.TUNE_KEYS[k], int(params[k])))

    def get_tune(self):
        out = {}
        for k, key in self.TUNE_KEYS.items():
            v = c_i64()
            check(lib().g5_msbfs_get_tune(self.h, key, ctypes.byref(v)))
            out[k] = v.value
        return out

    def autotune(self, roots, passes=1):
        r = np.ascontiguousarray(roots, dtype=np.int64)
        best = ctypes.c_double()
        trials = ctypes.c_int()
        check(lib().g5_msbfs_autotune(self.h, _p(r), r.size, passes, ctypes.byref(best), ctypes.byref(trials)))
        self.nroots = r.size
        return best.value, trials.value, self.get_tune()

    def place(self, roots, trials=4):
        """Re-allocates the parent records up to `trials` times and keeps the
        placement whose search of `roots` is fastest: (ms before, ms after,
        allocations kept).  The maps then hold the results of `roots`."""
        r = np.ascontiguousarray(roots, dtype=np.int64)
        first, best, moves = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        check(lib().g5_msbfs_place(self.h, _p(r), r.size, trials, ctypes.byref(first), ctypes.byref(best),
                                   ctypes.byref(moves)))
        self.nroots = r.size
        return first.value, best.value, moves.value

    # ---- continuous mode ----
    def open(self):
        check(lib().g5_msbfs_open(self.h))

    def join(self, roots, lanes):
        r = np.ascontiguousarray(roots, dtype=np.int64)
        ln = np.ascontiguousarray(lanes, dtype=np.int32)
        check(lib().g5_msbfs_join(self.h, _p(r), _p(ln), r.size))

    def step(self):
        live = c_u64()
        ms = ctypes.c_float()
        check(lib().g5_msbfs_step(self.h, ctypes.byref(live), ctypes.byref(ms)))
        return live.value, ms.value

    def leave(self, mask):
        ne = np.zeros(64, np.int64)
        ck = np.zeros(64, np.uint64)
        check(lib().g5_msbfs_leave(self.h, mask, _p(ne), _p(ck)))
        return ne, ck

    def nedge(self):
        out = np.empty(64, np.int64)
        check(lib().g5_msbfs_nedge(self.h, _p(out)))
        return out[:self.nroots].copy()

    def profile(self, enable=True):
        check(lib().g5_msbfs_profile(self.h, 1 if enable else 0))

    def stats(self):
        res = {}
        for k, name in enumerate(self.KERNELS):
            ms, bytes_ = ctypes.c_double(), ctypes.c_double()
            launches = c_i64()
            check(lib().g5_msbfs_stats(self.h, k, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(bytes_)))
            res[name] = {"ms": ms.value, "launches": launches.value, "bytes": bytes_.value}
        return res

    def levels(self):
        n = ctypes.c_int()
        check(lib().g5_msbfs_levels(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1) * 12, np.float64)
        check(lib().g5_msbfs_levels(self.h, _p(out), n.value, ctypes.byref(n)))
        return out[:n.value * 12].reshape(-1, 12)

    def set_mode(self, mode):
        """0: direction per level, 1: pull only, 2: push only."""
        check(lib().g5_msbfs_set_mode(self.h, mode))

    def bins(self):
        out = np.zeros(3, np.int64)
        check(lib().g5_msbfs_info(self.h, _p(out), 3))
        return out.tolist()

    def reach(self, roots):
        """Per root: the edges of its neighbours (at most 64 roots)."""
        r = np.ascontiguousarray(roots, dtype=np.int64)
        out = np.zeros(max(r.size, 1), np.int64)
        check(lib().g5_msbfs_reach(self.h, _p(r), r.size, _p(out)))
        return out[:r.size]

    def split_levels(self):
        """Levels since the last run/open that pulled some lanes and pushed the others."""
        out = np.zeros(4, np.int64)
        check(lib().g5_msbfs_info(self.h, _p(out), 4))
        return int(out[3])

    def free(self):
        if self.h:
            lib().g5_msbfs_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def s4_max_words(n):
    return lib().g5_s4_max_words(n)


def s4_encode(x):
    """Host uint32 array -> reference-identical word stream, encoded on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    din = DeviceArray.from_host(x) if x.size else DeviceArray(1, np.uint32)
    cap = s4_max_words(x.size)
    dout = DeviceArray(cap, np.uint32)
    nw = c_sz()
    check(lib().g5_s4_encode(din.ptr, x.size, dout.ptr, cap, ctypes.byref(nw)))
    return dout.to_host(nw.value)


def s4_decode(words, cap):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    din = DeviceArray.from_host(w)
    dout = DeviceArray(max(cap, 1), np.uint32)
    n = c_sz()
    check(lib().g5_s4_decode(din.ptr, w.size, dout.ptr, cap, ctypes.byref(n)))
    return dout.to_host(n.value)


def vreduce_or(maps):
    maps = np.ascontiguousarray(maps, dtype=np.uint64)
    k, n = maps.shape
    d = DeviceArray.from_host(maps.reshape(-1))
    o = DeviceArray(n, np.uint64)
    check(lib().g5_vreduce_or(d.ptr, k, n, o.ptr))
    return o.to_host()


def vreduce_min_pred(preds):
    preds = np.ascontiguousarray(preds, dtype=np.int64)
    k, n = preds.shape
    d = DeviceArray.from_host(preds.reshape(-1))
    o = DeviceArray(n, np.int64)
    check(lib().g5_vreduce_min_pred(d.ptr, k, n, o.ptr))
    return o.to_host()


def vreduce_or_list(maps):
    """The OR of k bitmaps held in k separate device buffers."""
    bufs = [DeviceArray.from_host(np.ascontiguousarray(m, dtype=np.uint64)) for m in maps]
    n = bufs[0].n
    ptrs = (ctypes.c_void_p * len(bufs))(*[b.ptr for b in bufs])
    o = DeviceArray(n, np.uint64)
    check(lib().g5_vreduce_or_list(ptrs, len(bufs), n, o.ptr))
    return o.to_host()


def vreduce_min_pred_list(preds):
    """The minimum present id over k predecessor arrays in separate buffers."""
    bufs = [DeviceArray.from_host(np.ascontiguousarray(p, dtype=np.int64)) for p in preds]
    n = bufs[0].n
    ptrs = (ctypes.c_void_p * len(bufs))(*[b.ptr for b in bufs])
    o = DeviceArray(n, np.int64)
    check(lib().g5_vreduce_min_pred_list(ptrs, len(bufs), n, o.ptr))
    return o.to_host()


BAD_FLAGS = {"range": 1, "root": 2, "self_parent": 4, "depth_map": 8, "edge_level": 16, "parent_edge": 32}


class Validator:
    """Graph500 result validation on the device (validate/validate.h:249-491)
    for a whole (1x1) graph.  validate() returns (bad, nedge, depth, ms):
    bad = 0 for a valid predecessor map, else the OR of BAD_FLAGS."""

    def __init__(self, graph):
        self.graph = graph
        self.h = c_vp()
        check(lib().g5_validator_create(graph.h, ctypes.byref(self.h)))

    def _call(self, fn, root, pred_ptr, nglobal, want_depth):
        bad, nedge, ms = ctypes.c_int(), c_i64(), ctypes.c_float()
        depth = np.empty(self.graph.n_rows, np.int32) if want_depth else None
        check(fn(self.h, root, self.graph.nglobal if nglobal is None else nglobal, pred_ptr, ctypes.byref(bad),
                 ctypes.byref(nedge), _p(depth), ctypes.byref(ms)))
        return bad.value, nedge.value, depth, ms.value

    def validate(self, root, pred, nglobal=None, want_depth=False):
        pred = np.ascontiguousarray(pred, dtype=np.int64)
        if pred.size != self.graph.n_rows:
            raise ValueError("pred must hold n_rows entries")
        return self._call(lib().g5_validate, root, _p(pred), nglobal, want_depth)

    def validate_device(self, root, d_pred, nglobal=None, want_depth=False):
        return self._call(lib().g5_validate_device, root, d_pred.ptr, nglobal, want_depth)

    def free(self):
        if self.h:
            lib().g5_validator_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bitmap_count(d_bits, n_words):
    """Set bits of a device bitmap (DeviceArray of uint64)."""
    c = c_i64()
    check(lib().g5_bitmap_count(d_bits.ptr, n_words, ctypes.byref(c)))
    return c.value


def s4_roundtrip(x, words=False):
    """Host uint32 array -> encoded and decoded on the GPU in one call;
    returns (decoded values, stream length in words), and the stream itself
    as a third item when words is true."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    din = DeviceArray.from_host(x)
    cap = s4_max_words(x.size)
    dw = DeviceArray(cap, np.uint32)
    dout = DeviceArray(max(x.size, 1), np.uint32)
    nw = c_sz()
    check(lib().g5_s4_roundtrip(din.ptr, x.size, dw.ptr, cap, dout.ptr, ctypes.byref(nw)))
    if words:
        return dout.to_host(x.size), nw.value, dw.to_host(nw.value)
    return dout.to_host(x.size), nw.value
