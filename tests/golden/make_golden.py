"""Write the golden fixtures the oracle is pinned against.

Run from the repository root in a container that holds the reference
checkout:  python tests/golden/make_golden.py

* kron_s{5,6,7,8}.npy     edge lists from the reference Graph500 generator
                          (cpu_2d/generator), compiled from its own sources by
                          oracle/Makefile into oracle/_ref/libref_gen.so
* kron_hashes.json        sha256 of the reference edge lists at larger scales
* codec_s4bp128d4.npz     input vectors and the word streams the reference
                          codec (s4-bp128-d4, cpu_2d/compression/cpusimd)
                          writes for them, via oracle/_ref/libref_codec.so
* roots_draws.json        first draws of std::uniform_int_distribution driven
                          by std::knuth_b (roots_draws.cpp, system g++)
All files are data (inputs and expected outputs); manifest.json records
provenance and sha256 of each.
"""
import ctypes
import hashlib
import json
import os
import subprocess
import tempfile
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/cpu_2d"


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def aligned_empty(n, dtype=np.uint32):
    raw = np.zeros(n + 16, dtype=dtype)
    off = (-raw.ctypes.data % 16) // raw.itemsize
    return raw[off:off + n]


def ref_generate(lib, scale, m):
    e = np.zeros(2 * m, np.int64)
    rc = lib.ref_generate(ctypes.c_int(scale), ctypes.c_int64(0), ctypes.c_int64(m),
                          ctypes.c_uint64(1), ctypes.c_uint64(2), _ptr(e))
    assert rc == 0
    return e


def ref_encode(lib, x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = aligned_empty(2 * len(x) + 64)
    n = ctypes.c_size_t(len(out))
    rc = lib.ref_encode(b"s4-bp128-d4", _ptr(x), ctypes.c_size_t(len(x)), _ptr(out),
                        ctypes.byref(n))
    assert rc == 0
    return out[:n.value].copy()


def codec_cases(oracle):
    rng = np.random.default_rng(20161213)
    cases = {}
    for n in (0, 1, 5, 127, 128, 129, 2047, 2048, 2049, 4096 + 77):
        cases[f"sorted_{n}"] = np.sort(rng.choice(1 << 22, size=n, replace=False)).astype(np.uint32)
        cases[f"random_{n}"] = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    cases["testsmall"] = np.array([1, 3, 5, 1039, 1045], np.uint32)  # src/unit.cpp:35-39
    cases["b0_const"] = np.full(256 + 9, 77, np.uint32)              # b = 0 miniblocks
    cases["b32_wide"] = np.sort(rng.integers(0, 1 << 32, size=384, dtype=np.uint64)).astype(np.uint32)
    cases["dense_run"] = np.arange(5000, 5000 + 3000, dtype=np.uint32)
    cases["max_values"] = np.full(130, 0xFFFFFFFF, np.uint32)
    # real BFS frontier queues: scale-16 Kronecker graph, first selected root
    m = 16 << 16
    seed = np.zeros(5, np.uint32)
    oracle.orc_make_mrg_seed(ctypes.c_uint64(1), ctypes.c_uint64(2), _ptr(seed))
    e = np.zeros(2 * m, np.int64)
    oracle.orc_generate_kronecker_range(_ptr(seed), 16, ctypes.c_int64(0), ctypes.c_int64(m), _ptr(e))
    oracle.orc_num_vertices.restype = ctypes.c_int64
    oracle.orc_padded_rows.restype = ctypes.c_int64
    oracle.orc_build_csr.restype = ctypes.c_int64
    nv = oracle.orc_num_vertices(_ptr(e), ctypes.c_int64(m))
    nr = oracle.orc_padded_rows(ctypes.c_int64(nv))
    rp = np.zeros(nr + 1, np.int64)
    cols = np.zeros(2 * m, np.uint32)
    oracle.orc_build_csr(_ptr(e), ctypes.c_int64(m), ctypes.c_int64(nr), _ptr(rp), _ptr(cols), None)
    roots = np.zeros(64, np.int64)
    oracle.orc_select_roots(_ptr(rp), ctypes.c_int64(nr), ctypes.c_int64(1 << 16), 64, _ptr(roots))
    pred = np.zeros(nr, np.int64)
    lev = np.zeros(nr, np.int32)
    nl = ctypes.c_int32()
    oracle.orc_bfs_bin(_ptr(rp), _ptr(cols), ctypes.c_int64(nr), ctypes.c_int64(int(roots[0])),
                       _ptr(pred), _ptr(lev), ctypes.byref(nl))
    for d in range(1, nl.value):
        cases[f"bfs_s16_level{d}"] = np.nonzero(lev == d)[0].astype(np.uint32)
    return cases


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle", "ref"])
    gen = ctypes.CDLL(os.path.join(ROOT, "oracle/_ref/libref_gen.so"))
    codec = ctypes.CDLL(os.path.join(ROOT, "oracle/_ref/libref_codec.so"))
    oracle = ctypes.CDLL(os.path.join(ROOT, "oracle/liboracle.so"))
    manifest = {"provenance": {
        "generator": f"{REF}/generator compiled by oracle/Makefile (ref target), seeds (1,2), edgefactor 16",
        "codec": f"{REF}/compression/cpusimd compiled by oracle/Makefile, CODECFactory 's4-bp128-d4', 16-byte aligned buffers",
        "roots": "tests/golden/roots_draws.cpp built with the system g++ (libstdc++ std::knuth_b)"},
        "files": {}}

    def record(name):
        with open(os.path.join(HERE, name), "rb") as f:
            manifest["files"][name] = hashlib.sha256(f.read()).hexdigest()

    for s in (5, 6, 7, 8):
        e = ref_generate(gen, s, 16 << s).reshape(-1, 2)
        np.save(os.path.join(HERE, f"kron_s{s}.npy"), e)
        record(f"kron_s{s}.npy")
    hashes = {}
    for s in (10, 12, 16, 18):
        e = ref_generate(gen, s, 16 << s)
        hashes[str(s)] = {"m": 16 << s, "sha256": hashlib.sha256(e.tobytes()).hexdigest(),
                          "sum": int(e.sum()), "xor": int(np.bitwise_xor.reduce(e))}
    with open(os.path.join(HERE, "kron_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1)
    record("kron_hashes.json")

    cases = codec_cases(oracle)
    arrays = {}
    for k, x in cases.items():
        arrays["in_" + k] = x
        arrays["out_" + k] = ref_encode(codec, x)
    np.savez_compressed(os.path.join(HERE, "codec_s4bp128d4.npz"), **arrays)
    record("codec_s4bp128d4.npz")

    draws = {}
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "roots_draws")
        subprocess.check_call(["g++", "-O2", "-std=c++11", "-o", exe, os.path.join(HERE, "roots_draws.cpp")])
        for s in (5, 10, 16, 20, 22, 24, 26):
            out = subprocess.check_output([exe, str(1 << s), "4096"]).decode().split()
            draws[str(s)] = [int(v) for v in out]
    with open(os.path.join(HERE, "roots_draws.json"), "w") as f:
        json.dump(draws, f)
    record("roots_draws.json")

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote", len(manifest["files"]), "fixtures")


if __name__ == "__main__":
    sys.exit(main())
