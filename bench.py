#!/usr/bin/env python3
"""Headline benchmark: Graph500 BFS on one MI355X (gfx950).

Workload (scale 26 by default): the Graph500 Kronecker graph (A=.57, B=C=.19,
edgefactor 16, seeds 1 and 2) generated and built into the 1x1 CSR block on
the device; 64 search keys chosen as cpu_2d/main.cpp:351-411 does.

* single-root: every root searched on its own (g5_bfs_run, resident in HBM);
  TEPS_r = nedge_r / t_r, t_r = device time of that search.
* batched: the 64 roots searched together by the multi-source kernel
  (g5_msbfs_run); TEPS_r = nedge_r / (T_batch / 64).
A step is one batched search over the 64 roots.  The reported value is the
harmonic mean of the batched TEPS_r over the roots (cpu_2d/main.cpp:944-950),
in GTEPS; the single-root harmonic mean and the ratio are reported beside it.

    python bench.py [--steps K] [--warmup W] [--scale S] [--roots R]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS hmean, batched-64 vs single-root, scale 26, 1 MI355X; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def hmean(x):
    x = np.asarray(x, dtype=np.float64)
    return float(len(x) / np.sum(1.0 / x))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(g, scale, root, nedge_gpu):
    """The oracle's CPUBFS_bin restatement (single thread) on this host, one
    root of the same graph: the bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    rp, cols, gdeg = g.download()

    class _G:
        pass

    ref = _G()
    ref.rowptr, ref.cols, ref.n_rows, ref.gdeg = rp, cols, g.n_rows, gdeg
    t0 = time.perf_counter()
    pred, lev, nl = O.bfs_bin(ref, int(root))
    t = time.perf_counter() - t0
    ne = O.nedge(ref, lev)
    return {
        "value": ne / t / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"1 root ({int(root)}) of the scale-{scale} graph, full search, {t:.1f} s, "
                  f"single thread on {cpu_model()} (nproc {os.cpu_count()})",
        "nedge_matches_gpu": int(ne) == int(nedge_gpu),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--levels", action="store_true", help="print per-level kernel accounting to stderr")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="WORDS",
                    help="only run the plain copy kernel (k_or over one map) on WORDS uint64 words, "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the 2D row/column exchange "
                                   "is emulated on a single device (see DESIGN.md)",
                          "n_gpus": args.gpus}))
        return 2

    import gpugraph500_amd as g5

    g5.init(0)
    if args.copy_calibration:
        n = args.copy_calibration
        src = g5.DeviceArray(n, np.uint64)
        dst = g5.DeviceArray(n, np.uint64)
        for _ in range(3):
            g5.check(g5.lib().g5_vreduce_or(src.ptr, 1, n, dst.ptr))
        print(json.dumps({"copy_calibration_words": n, "bytes_read": 8 * n, "bytes_written": 8 * n}))
        return 0
    t0 = time.perf_counter()
    g = g5.Graph.kronecker(args.scale)
    log(f"graph: scale {args.scale}, {g.nglobal} vertices, {g.nnz} nonzeros, "
        f"built on device in {time.perf_counter() - t0:.2f} s")
    roots = g.select_roots(1 << args.scale, args.roots)
    log(f"roots: {len(roots)}")

    # ---- single-root searches (each root once after warmup) ----
    b = g5.BFS(g)
    for r in roots[:max(args.warmup, 1)].tolist():
        b.run(r)
    b.profile(True)
    single_ms, nedge = [], []
    for r in roots.tolist():
        nl, ms = b.run(r)
        single_ms.append(ms)
        nedge.append(b.nedge())
    st = b.stats()
    nedge = np.array(nedge, np.float64)
    single_teps = nedge / (np.array(single_ms) / 1e3)
    single_h = hmean(single_teps) / 1e9
    log(f"single-root: hmean {single_h:.3f} GTEPS, mean {np.mean(single_ms):.2f} ms/root")

    expand_gbs = st["bytes"] / (st["ms"] / 1e3) / 1e9 if st["ms"] > 0 else 0.0
    single_kernel = {"kernel": "k_expand", "achieved_gbs": expand_gbs,
                     "avg_launch_ms": st["ms"] / max(st["launches"], 1),
                     "algorithmic_bytes_per_launch": st["bytes"] / max(st["launches"], 1)}
    b.free()

    # ---- batched multi-source search over the same roots ----
    m = g5.MSBFS(g)
    log(f"msbfs degree bins (>2048, 33..2048, 1..32): {m.bins()}")
    for _ in range(max(args.warmup, 1)):
        m.run(roots)
    m.profile(True)            # one run as direct launches with per-kernel events
    m.run(roots)
    kst = m.stats()
    if args.levels:
        for i, rec in enumerate(m.levels()):
            log(f"level {i + 1} {'push' if rec[9] else 'pull'}: ms " + "/".join(f"{x:.2f}" for x in rec[:4]) +
                " edges " + "/".join(f"{x:.2e}" for x in rec[4:8]) +
                f" new {rec[8]:.3e} frontier {rec[10]:.3e} v {rec[11]:.3e} e")
    m.profile(False)
    dom = max(kst, key=lambda k: kst[k]["ms"])
    d = kst[dom]
    dom_gbs = d["bytes"] / (d["ms"] / 1e3) / 1e9 if d["ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dom_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "avg_launch_ms": d["ms"] / max(d["launches"], 1),
                "algorithmic_bytes_per_launch": d["bytes"] / max(d["launches"], 1),
                "all_kernels": {k: {"ms": v["ms"], "launches": v["launches"],
                                    "gbs": (v["bytes"] / (v["ms"] / 1e3) / 1e9) if v["ms"] > 0 else 0.0}
                                for k, v in kst.items()}}

    # ---- timed steps: one batched search of all roots per step ----
    step_ms, dev_ms = [], []
    for i in range(args.warmup + args.steps):
        g5.lib().g5_device_sync()
        t = time.perf_counter()
        nl, ms = m.run(roots)
        g5.lib().g5_device_sync()
        if i >= args.warmup:
            step_ms.append((time.perf_counter() - t) * 1e3)
            dev_ms.append(ms)
    ne_b = m.nedge().astype(np.float64)
    assert np.array_equal(ne_b, nedge), "batched and single-root TEPS numerators differ"
    t_batch = float(np.mean(dev_ms)) / 1e3
    batched_teps = ne_b / (t_batch / len(roots))
    value = hmean(batched_teps) / 1e9
    aggregate = float(ne_b.sum() / t_batch / 1e9)
    log(f"batched: {t_batch * 1e3:.1f} ms per 64-root search, hmean {value:.2f} GTEPS, "
        f"aggregate {aggregate:.2f} GTEPS, {nl} levels")
    mode = "batched-64"

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids / u64 bitmaps (integer)",
        "data": "synthetic: Graph500 Kronecker (A=.57 B=C=.19), edgefactor 16, seeds (1,2), generated on the device",
        "config": {"workload": f"{mode} BFS, 64 roots, Kronecker scale {args.scale}, 1x1 block",
                   "scale": args.scale, "edgefactor": 16, "roots": int(len(roots)),
                   "vertices": int(g.nglobal), "nnz": int(g.nnz)},
        "single_root_hmean_gteps": single_h,
        "batched_hmean_gteps": value,
        "batched_over_single": value / single_h,
        "aggregate_gteps": aggregate,
        "batched_ms": t_batch * 1e3,
        "single_root_mean_ms": float(np.mean(single_ms)),
        "single_root_kernel": single_kernel,
        "roofline": roofline,
    }
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(g, args.scale, roots[0], nedge[0])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
