#!/usr/bin/env python3
"""Headline benchmark: Graph500 BFS on one MI355X (gfx950).

Workload: the Graph500 Kronecker graph (A=.57, B=C=.19, edgefactor 16, seeds
1 and 2) generated and built on the device; 64 search keys chosen as
cpu_2d/main.cpp:351-411 does.  TEPS_r = nedge_r / t_r with nedge from
validate/validate.h:439-485; the reported value is the harmonic mean over
roots (cpu_2d/main.cpp:944-950), in GTEPS.

--config c4 (default), scale 26, 1x1 block
    A step is one batched search of the 64 roots (g5_msbfs_run), inputs
    resident in HBM; batched TEPS_r = nedge_r / (T_batch / 64).  The plain
    single-root search (g5_bfs_run) of every root is measured beside it.
--config c2, scale 22
    A step is the 64 single-root searches with every level's queue through
    s4-bp128-d4 encode -> decode on the device (G5_RUN_CODEC).
--config c3, scale 26
    One rank's share of a 4 x 4 grid: the g500_hip driver's emulated grid
    (device exchange, codec on the fold), rank (0, 0)'s share timed.
--config c1, scale 16
    The CPU-runnable case: a step is the 64 single-root searches on the
    GPU; the CPU reference algorithm (the oracle's CPUBFS_bin restatement,
    one thread) runs every one of the 64 roots beside it, and the batched
    search of the same roots is reported too.

The batched time runs from the start to the 64 final predecessor and level
maps in HBM (parent int32[n][64], level uint8[n][64], original ids), as the
reference's runBFS window ends with the predecessors final
(cpu_2d/main.cpp:419-462); the single-root time likewise ends with its
parent and level arrays final.  Every measured search is then validated on
the device (g5_validate_device: the checks of validate/validate.h:249-491
with its TEPS numerator), and every batched lane's predecessor and level
maps are compared word for word, on the device, with the single-root search
of the same root; a failure aborts the run.

Shapes: the timed roots are the first 64 of the benchmark's draw; the
batched search's shapes come from tuned_shapes.json for this graph's scale
and degree histogram, or from an autotune on the next 64 draws (disjoint
from the timed set), which is then persisted.

roofline.traffic: HBM bytes per launch of the dominant kernel from two
rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) run as child processes before
this process touches the GPU, each scaled by what the counter reads for a
plain copy kernel of known size; null when rocprofv3 is unavailable.  A
first child (no profiler) finds the shapes; the counter passes and the timed
run all use those shapes, so the kernels measured are the kernels timed.

    python bench.py [--config c4|c2|c3|c1] [--steps K] [--warmup W] [--scale S]
"""
import argparse
import csv
import glob
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS hmean, batched-64 vs single-root, scale 26, 1 MI355X; HBM GB/s"
HBM_PEAK_GBS = 8000.0
COPY_WORDS = 1 << 26          # copy-kernel calibration: 64 Mi uint64 words read and written
TUNE_PASSES = 2               # coordinate-descent rounds of the batched search's autotune
PULL_KERNELS = ("k_pull_block", "k_pull_wave", "k_pull_thread")
DEFAULT_SCALE = {"c4": 26, "c2": 22, "c3": 26, "c1": 16}


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


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def _host_csr(g):
    rp, cols, gdeg = g.download()

    class _G:
        pass

    ref = _G()
    ref.rowptr, ref.cols, ref.n_rows, ref.gdeg = rp, cols, g.n_rows, gdeg
    return ref


def cpu_baseline(g, scale, roots, nedge_gpu, budget_s=10.0):
    """The oracle's CPUBFS_bin restatement (single thread) on this host: whole
    searches of the same graph's roots, in order, until about budget_s of
    CPU work (at least one) -- the bounded sample of the workload; TEPS is
    the harmonic mean over the searches done."""
    O = _oracle()
    ref = _host_csr(g)
    teps, done, t_all, same = [], 0, 0.0, True
    for k, root in enumerate(roots):
        t0 = time.perf_counter()
        pred, lev, nl = O.bfs_bin(ref, int(root))
        t = time.perf_counter() - t0
        ne = O.nedge(ref, lev)
        same = same and int(ne) == int(nedge_gpu[k])
        teps.append(ne / t)
        done += 1
        t_all += t
        if t_all >= budget_s:
            break
    return {
        "value": hmean(teps) / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of {len(roots)} roots of the scale-{scale} graph, whole searches, {t_all:.1f} s, "
                  f"single thread on {cpu_model()} (nproc {os.cpu_count()})",
        "nedge_matches_gpu": same,
    }


def cpu_baseline_codec(g, scale, roots, nedge_gpu, budget_s=12.0):
    """Single-root search plus s4-bp128-d4 encode -> decode of every level's
    sorted queue, both from the oracle, single thread, on as many roots as
    fit the budget."""
    O = _oracle()
    ref = _host_csr(g)
    teps, done, t_all = [], 0, 0.0
    for k, root in enumerate(roots):
        t0 = time.perf_counter()
        pred, lev, nl = O.bfs_bin(ref, int(root))
        for d in range(1, nl + 1):
            q = np.flatnonzero(lev == d).astype(np.uint32)
            if q.size:
                w = O.s4_encode(q)
                rc, back = O.s4_decode(w, q.size)
                assert rc == 0 and np.array_equal(back, q)
        t = time.perf_counter() - t0
        ne = O.nedge(ref, lev)
        assert int(ne) == int(nedge_gpu[k])
        teps.append(ne / t)
        done += 1
        t_all += t
        if t_all > budget_s:
            break
    return {
        "value": hmean(teps) / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of {len(roots)} roots of the scale-{scale} graph, search + codec round trip of every "
                  f"level queue, {t_all:.1f} s, single thread on {cpu_model()} (nproc {os.cpu_count()})",
    }


# ---------------------------------------------------------------- PMC traffic

def root_sets(g, scale, n):
    """The timed search keys (the first n of the benchmark's draw,
    cpu_2d/main.cpp:351-411) and a disjoint set for tuning (the next n draws
    of the same generator); the timed set again when the graph has too few
    eligible vertices."""
    allr = g.select_roots(1 << scale, 2 * n)
    timed = allr[:n]
    tune = allr[n:2 * n] if len(allr) >= 2 * n else timed
    return timed, tune


def tune_probe(args):
    """Child process (no profiler): the batched search's shapes for this
    graph -- the persisted ones for its scale and degree histogram, or an
    autotune on the tuning root set (disjoint from the timed set), then
    persisted -- printed as one JSON line."""
    import gpugraph500_amd as g5
    from gpugraph500_amd import shapes
    g5.init(0)
    g = g5.Graph.kronecker(args.scale)
    timed, tune = root_sets(g, args.scale, args.roots)
    key = shapes.graph_key(g, args.scale)
    saved = shapes.load(key, g5.MSBFS.TUNE_KEYS)
    if saved is not None:
        print(json.dumps({"params": saved["params"], "candidates_run": 0, "best_ms": saved.get("best_ms"),
                          "source": f"persisted ({shapes.PATH}, key {key})", "tuned_on": saved.get("tuned_on")}),
              flush=True)
        return 0
    m = g5.MSBFS(g)
    m.run(tune)
    # tune in the fast record placement, the one the timed searches run in
    # (the timed run places its records the same way before it times)
    m.place(tune, 4)
    best_ms, trials, params = m.autotune(tune, passes=TUNE_PASSES)
    shapes.save(key, {"params": params, "best_ms": best_ms, "tuned_on": "roots 65-128 of the benchmark's draw"})
    print(json.dumps({"params": params, "candidates_run": trials, "best_ms": best_ms,
                      "source": f"autotuned now, persisted under key {key}",
                      "tuned_on": "roots 65-128 of the benchmark's draw"}), flush=True)
    return 0


def pmc_probe(args):
    """Run under rocprofv3 --pmc by measure_traffic: two batched searches with
    the given shapes and the copy calibration kernel."""
    import gpugraph500_amd as g5
    g5.init(0)
    g = g5.Graph.kronecker(args.scale)
    roots = g.select_roots(1 << args.scale, args.roots)
    m = g5.MSBFS(g)
    if args.tune_json:
        m.tune(**json.loads(args.tune_json))
    m.run(roots)
    m.run(roots)
    m.free()
    g.free()
    src = g5.DeviceArray(COPY_WORDS, np.uint64)
    dst = g5.DeviceArray(COPY_WORDS, np.uint64)
    for _ in range(3):
        g5.check(g5.lib().g5_vreduce_or(src.ptr, 1, COPY_WORDS, dst.ptr))
    g5.check(g5.lib().g5_device_sync())
    log("pmc probe done")
    return 0


def probe(args):
    """Workloads for profiler runs (rocprofv3 before this script):
    single    the 64 single-root searches (profiled) and the copy calibration kernel
    exchange  the c3 and c2 sub-records by themselves
    batched   one profiled and six timed batched searches with the given shapes
              (after the parent records' placement, --place trials)
    Each prints one JSON line."""
    import gpugraph500_amd as g5
    g5.init(0)
    g = g5.Graph.kronecker(args.scale)
    roots, tune_roots = root_sets(g, args.scale, args.roots)
    if args.probe == "single":
        b = g5.BFS(g)
        for r in roots[:2].tolist():
            b.run(r)
        b.profile(True)
        ms = [b.run(r)[1] for r in roots.tolist()]
        out = {"mean_ms": float(np.mean(ms)), "kernels": b.kernel_stats()}
        b.free()
        g.free()
        src = g5.DeviceArray(COPY_WORDS, np.uint64)
        dst = g5.DeviceArray(COPY_WORDS, np.uint64)
        for _ in range(3):
            g5.check(g5.lib().g5_vreduce_or(src.ptr, 1, COPY_WORDS, dst.ptr))
        g5.check(g5.lib().g5_device_sync())
    elif args.probe == "exchange":
        out = {"c3": sub_c3(g5, g, roots)}
        g.free()
        out["c2"] = sub_c2(g5)
    else:
        m = g5.MSBFS(g)
        if args.tune_json:
            m.tune(**json.loads(args.tune_json))
        m.run(roots)
        if args.place > 0:
            m.place(tune_roots, args.place)
        m.profile(True)
        m.run(roots)
        st, lv = m.stats(), m.levels()
        m.profile(False)
        ts = [m.run(roots)[1] for _ in range(6)]
        out = {"ms": float(np.mean(ts[1:])), "times": m.times(), "kernels": st,
               "levels": [[round(x, 3) for x in r[:4]] + [int(r[9])] for r in lv]}
    print(json.dumps(out), flush=True)
    return 0


def _pmc_pass(exe, counter, args, tune_json):
    """{pull kernel: mean counter per dispatch}, the copy kernel's mean and the
    dispatches sampled."""
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        cmd = [exe, "--pmc", counter, "--kernel-include-regex", "k_pull|k_or", "-d", d, "-o", "p",
               "--output-format", "csv", "--", sys.executable, os.path.abspath(__file__), "--pmc-probe",
               "--scale", str(args.scale), "--roots", str(args.roots), "--tune-json", tune_json]
        env = dict(os.environ, TMPDIR="/tmp")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"rocprofv3 --pmc {counter} exited {r.returncode}: {r.stdout[-400:]!r}")
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file")
        kern, copy = {k: [] for k in PULL_KERNELS}, []
        with open(files[0]) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                name, v = row["Kernel_Name"], float(row["Counter_Value"])
                for k in PULL_KERNELS:
                    if k in name:
                        kern[k].append(v)
                if "k_or" in name:
                    copy.append(v)
        if not copy or not any(kern.values()):
            raise RuntimeError(f"rocprofv3 --pmc {counter}: kernels not found")
        return ({k: float(np.mean(v)) for k, v in kern.items() if v}, float(np.mean(copy)),
                {k: len(v) for k, v in kern.items()})


def run_tune_probe(args):
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--tune-probe", "--scale", str(args.scale),
                        "--roots", str(args.roots)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"tune probe exited {r.returncode}: {r.stderr[-400:]!r}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def measure_traffic(args):
    """Shapes from a tune probe, then HBM bytes per launch of every pull
    kernel with those shapes from FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled so that the copy kernel reads and writes
    8 * COPY_WORDS bytes.  Returns ({kernel: bytes}, info, tuned)."""
    t0 = time.perf_counter()
    try:
        tuned = run_tune_probe(args) if args.autotune == "on" else None
    except (RuntimeError, subprocess.TimeoutExpired, OSError, ValueError) as e:
        return None, {"error": str(e)[:300]}, None
    if tuned:
        tuned["seconds"] = time.perf_counter() - t0
        log(f"tune probe: {tuned['params']} after {tuned['candidates_run']} candidates, {tuned['best_ms']:.2f} ms")
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, {"error": "rocprofv3 not found"}, tuned
    tune_json = json.dumps(tuned["params"]) if tuned else ""
    try:
        t1 = time.perf_counter()
        f_k, f_c, n_f = _pmc_pass(exe, "FETCH_SIZE", args, tune_json)
        w_k, w_c, n_w = _pmc_pass(exe, "WRITE_SIZE", args, tune_json)
    except (RuntimeError, subprocess.TimeoutExpired, OSError) as e:
        return None, {"error": str(e)[:300]}, tuned
    copy_bytes = 8.0 * COPY_WORDS
    f_scale = copy_bytes / (f_c * 1024.0)
    w_scale = copy_bytes / (w_c * 1024.0)
    traffic = {k: f_k[k] * 1024.0 * f_scale + w_k.get(k, 0.0) * 1024.0 * w_scale for k in f_k}
    info = {"fetch_kib_per_launch": f_k, "write_kib_per_launch": w_k, "fetch_scale": f_scale,
            "write_scale": w_scale, "launches_sampled": n_f, "shapes": tuned["params"] if tuned else "default",
            "seconds": time.perf_counter() - t1}
    log("traffic per launch: " + ", ".join(f"{k} {v / 1e9:.2f} GB" for k, v in traffic.items()) +
        f" (fetch x{f_scale:.2f}, write x{w_scale:.2f})")
    return traffic, info, tuned


# ---------------------------------------------------------------- configs

def bench_c4(args, g5, traffic, traffic_info, probe_tuned):
    t0 = time.perf_counter()
    g = g5.Graph.kronecker(args.scale)
    log(f"graph: scale {args.scale}, {g.nglobal} vertices, {g.nnz} nonzeros, "
        f"built on device in {time.perf_counter() - t0:.2f} s")
    roots, tune_roots = root_sets(g, args.scale, args.roots)
    log(f"roots: {len(roots)} timed, {len(tune_roots)} for tuning (disjoint: {len(set(roots) & set(tune_roots)) == 0})")

    # ---- single-root searches (each root once after warmup) ----
    b = g5.BFS(g)
    for r in roots[:max(args.warmup, 1)].tolist():
        b.run(r)
    b.profile(True)
    single_ms, nedge = [], []
    val = g5.Validator(g) if args.validate == "on" else None
    d_pred = g5.DeviceArray(g.n_rows, np.int64) if val else None
    val_ms, val_fail = [], 0
    for r in roots.tolist():
        nl, ms = b.run(r)
        single_ms.append(ms)
        nedge.append(b.nedge())
        if val:  # outside the search's own device-time window
            b.pred_device(d_pred)
            bad, vne, _, vms = val.validate_device(r, d_pred)
            val_fail += int(bad != 0 or vne != nedge[-1])
            val_ms.append(vms)
    kst1 = b.kernel_stats()
    nedge = np.array(nedge, np.float64)
    single_teps = nedge / (np.array(single_ms) / 1e3)
    single_h = hmean(single_teps) / 1e9
    log(f"single-root: hmean {single_h:.3f} GTEPS, mean {np.mean(single_ms):.2f} ms/root")
    single_kernel = {k: {"achieved_gbs": (v["bytes"] / (v["ms"] / 1e3) / 1e9) if v["ms"] > 0 else 0.0,
                         "avg_launch_ms": v["ms"] / max(v["launches"], 1), "launches": v["launches"],
                         "algorithmic_bytes_per_launch": v["bytes"] / max(v["launches"], 1)}
                     for k, v in kst1.items()}
    b.free()

    # ---- batched multi-source search over the same roots ----
    m = g5.MSBFS(g)
    for _ in range(max(args.warmup, 1)):
        m.run(roots)
    tuned = None
    if probe_tuned:  # the shapes the counter passes measured
        m.tune(**probe_tuned["params"])
        tuned = dict(probe_tuned, probe="tune probe (child process, before the counter passes)")
        log(f"shapes from the tune probe: {probe_tuned['params']} ({probe_tuned.get('source')})")
    elif args.autotune == "on":
        t0 = time.perf_counter()
        if args.place > 0:
            m.place(tune_roots, args.place)
        best_ms, trials, params = m.autotune(tune_roots, passes=TUNE_PASSES)
        tuned = {"params": params, "candidates_run": trials, "best_ms": best_ms,
                 "seconds": time.perf_counter() - t0, "source": "this process",
                 "tuned_on": "roots 65-128 of the benchmark's draw"}
        log(f"autotune: {params} after {trials} candidates, {best_ms:.2f} ms")
    if tuned is not None:
        tuned["timed_on"] = "roots 1-64 of the benchmark's draw"
        tuned["disjoint"] = bool(len(set(roots.tolist()) & set(tune_roots.tolist())) == 0)
    log(f"msbfs degree bins (workgroup, wavefront, thread per vertex): {m.bins()}")
    # the parent records' physical placement sets one of two speeds: keep the
    # faster of a few fresh allocations (setup, on the tuning roots)
    placement = None
    if args.place > 0:
        t0 = time.perf_counter()
        first, best, moves = m.place(tune_roots, args.place)
        placement = {"trials": args.place, "ms_before": first, "ms_after": best, "moves": moves,
                     "seconds": time.perf_counter() - t0, "on": "roots 65-128 of the benchmark's draw",
                     "what": "g5_msbfs_place: fresh allocations of the parent records (copies), the one with "
                             "the fastest search kept"}
        log(f"record placement: {first:.2f} -> {best:.2f} ms ({moves} moves in {args.place} trials)")
    m.profile(True)            # one run as direct launches with per-kernel events
    m.run(roots)
    kst = m.stats()
    if args.levels:
        for i, rec in enumerate(m.levels()):
            kind = {0: "pull", 1: "push", 2: "split"}[int(rec[9])]
            log(f"level {i + 1} {kind}: ms " + "/".join(f"{x:.2f}" for x in rec[:4]) +
                " edges " + "/".join(f"{x:.2e}" for x in rec[4:8]) +
                f" new {rec[8]:.3e} frontier {rec[10]:.3e} v {rec[11]:.3e} e")
    m.profile(False)
    dom = max(kst, key=lambda k: kst[k]["ms"])
    d = kst[dom]
    dom_gbs = d["bytes"] / (d["ms"] / 1e3) / 1e9 if d["ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic.get(dom) if traffic else None,
                "kernel": dom, "avg_launch_ms": d["ms"] / max(d["launches"], 1),
                "algorithmic_bytes_per_launch": d["bytes"] / max(d["launches"], 1),
                "traffic_source": traffic_info,
                "traffic_all_kernels": traffic,
                # the same kernel's measured HBM bytes per launch over its
                # launch time: how close its line fills come to the peak
                "traffic_gbs": (traffic[dom] / (d["ms"] / max(d["launches"], 1) / 1e3) / 1e9
                                if traffic and dom in traffic and d["ms"] > 0 else None),
                "all_kernels": {k: {"ms": v["ms"], "launches": v["launches"],
                                    "gbs": (v["bytes"] / (v["ms"] / 1e3) / 1e9) if v["ms"] > 0 else 0.0}
                                for k, v in kst.items()}}

    # ---- timed steps: one batched search of all roots per step, from the
    # start to the 64 final predecessor and level maps in HBM ----
    step_ms, dev_ms, search_ms, fin_ms = [], [], [], []
    for i in range(args.warmup + args.steps):
        g5.lib().g5_device_sync()
        t = time.perf_counter()
        nl, ms = m.run(roots)
        g5.lib().g5_device_sync()
        if i >= args.warmup:
            step_ms.append((time.perf_counter() - t) * 1e3)
            dev_ms.append(ms)
            a, b_ = m.times()
            search_ms.append(a)
            fin_ms.append(b_)
    levels_final = m.maps()[3]
    ne_b = m.nedge().astype(np.float64)
    assert np.array_equal(ne_b, nedge), "batched and single-root TEPS numerators differ"
    validation = None
    if val:
        for lane, r in enumerate(roots.tolist()):
            m.pred_device(lane, d_pred)
            bad, vne, _, vms = val.validate_device(r, d_pred)
            val_fail += int(bad != 0 or vne != ne_b[lane])
            val_ms.append(vms)
        validation = {"searches": len(val_ms), "failures": val_fail, "mean_ms": float(np.mean(val_ms)),
                      "what": "every single-root and batched result, device validator (validate.h:249-491)"}
        log(f"validation: {len(val_ms)} results, {val_fail} failures, {np.mean(val_ms):.1f} ms each")
        # bit-exact agreement of every batched lane with the single-root
        # search of the same root: predecessor and level maps, compared on
        # the device
        b2 = g5.BFS(g)
        d_single = g5.DeviceArray(g.n_rows, np.int64)
        l_single, l_batch = g5.DeviceArray(g.n_rows, np.int32), g5.DeviceArray(g.n_rows, np.int32)
        differ = []
        for lane, r in enumerate(roots.tolist()):
            b2.run(r)
            b2.pred_device(d_single)
            b2.level_device(l_single)
            m.pred_device(lane, d_pred)
            m.level_device(lane, l_batch)
            if g5.diff_u32(d_single, d_pred, 2 * g.n_rows) or g5.diff_u32(l_single, l_batch, g.n_rows):
                differ.append(lane)
        for a in (b2, d_single, l_single, l_batch):
            a.free()
        validation["batched_equals_single_root"] = {
            "lanes_compared": len(roots), "lanes_differing": differ,
            "what": "every lane's predecessor and level maps against the single-root search of its root, "
                    "compared word for word on the device"}
        log(f"batched == single-root predecessor and level maps: {len(roots) - len(differ)} of {len(roots)} lanes")
        val.free()
        d_pred.free()
        if val_fail:
            raise SystemExit(f"validation failed for {val_fail} of {len(val_ms)} results")
        if differ:
            raise SystemExit(f"batched and single-root maps differ on lanes {differ}")
    t_batch = float(np.mean(dev_ms)) / 1e3
    batched_teps = ne_b / (t_batch / len(roots))
    value = hmean(batched_teps) / 1e9
    aggregate = float(ne_b.sum() / t_batch / 1e9)
    log(f"batched: {t_batch * 1e3:.1f} ms per 64-root search, hmean {value:.2f} GTEPS, "
        f"aggregate {aggregate:.2f} GTEPS, {nl} levels")
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
        "config": {"workload": f"batched-64 BFS, 64 roots, Kronecker scale {args.scale}, 1x1 block",
                   "scale": args.scale, "edgefactor": 16, "roots": int(len(roots)),
                   "vertices": int(g.nglobal), "nnz": int(g.nnz)},
        "single_root_hmean_gteps": single_h,
        "batched_hmean_gteps": value,
        "batched_over_single": value / single_h,
        "aggregate_gteps": aggregate,
        "batched_ms": t_batch * 1e3,
        "batched_ms_split": {"search": float(np.mean(search_ms)), "finalize": float(np.mean(fin_ms)),
                             "what": "search: start to the last level (the kernels that reach a pair write its "
                                     "parent, as an original id at the vertex's original index, and its level); "
                                     "finalize: the pass that marks the used lanes' unreached slots"},
        "batched_output": {"layout": "parent int32[n][64] and level uint8[n][64] by original vertex id, "
                                     "slot = lane, in HBM (g5_msbfs_maps)",
                           "levels_final": bool(levels_final)},
        "single_root_mean_ms": float(np.mean(single_ms)),
        "single_root_kernel": single_kernel,
        "roofline": roofline,
        "autotune": tuned,
        "placement": placement,
        "validation": validation,
    }
    m.free()
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(g, args.scale, roots, nedge)
    if args.subconfigs == "on":
        out["c3"] = sub_c3(g5, g, roots)
        g.free()
        out["c2"] = sub_c2(g5)
    return out


def _rates(stats):
    """GB/s per exchange unit from g5_exchange_stats (algorithmic bytes over
    the units' device time)."""
    return {k: {"gbs": (v["bytes"] / (v["ms"] / 1e3) / 1e9) if v["ms"] > 0 else 0.0,
                "frac": (v["bytes"] / (v["ms"] / 1e3) / 1e9) / HBM_PEAK_GBS if v["ms"] > 0 else 0.0,
                "ms": v["ms"], "calls": v["calls"], "bytes": v["bytes"]}
            for k, v in stats.items() if v["calls"]}


def sub_c2(g5, scale=22, nroots=64):
    """C2 beside the headline: the single-root search with every level's queue
    through s4-bp128-d4 encode -> decode on the device at scale 22 (hmean
    GTEPS over the 64 search keys, each checked against the plain search's
    TEPS numerator), and the device codec's own rates, event-timed through
    g5_s4_encode / g5_s4_decode on the level queues of four of those
    searches (encode 4n + 4*csize bytes, decode 4*csize + 4n)."""
    t0 = time.perf_counter()
    g = g5.Graph.kronecker(scale)
    roots = g.select_roots(1 << scale, nroots)
    b = g5.BFS(g)
    b.run(int(roots[0]), codec=True)
    ms, teps = [], []
    for r in roots.tolist():
        b.run(r)
        ne = b.nedge()
        nl, t = b.run(r, codec=True)
        assert b.nedge() == ne, "the codec changed a search"
        ms.append(t)
        teps.append(ne / (t / 1e3))
    queues = []
    for r in roots[:4].tolist():
        b.run(r)
        _, lev = b.result()
        queues += [np.flatnonzero(lev == d).astype(np.uint32) for d in range(1, int(lev.max()) + 1)]
    b.free()
    g.free()
    words, sizes = 0, 0
    # code objects loaded before timing: the long and the short streams' kernels
    for warm in (np.arange(0, 1 << 20, 3, dtype=np.uint32), np.arange(0, 3000, 3, dtype=np.uint32)):
        g5.s4_decode(g5.s4_encode(warm), warm.size)
    g5.exchange_profile(True)
    for q in queues:
        w = g5.s4_encode(q)
        back = g5.s4_decode(w, q.size)
        assert np.array_equal(back, q), "codec round trip lost values"
        words += w.size
        sizes += q.size
    rates = _rates(g5.exchange_stats())
    g5.exchange_profile(False)
    log(f"c2: hmean {hmean(teps) / 1e9:.2f} GTEPS with the codec on every level; codec encode "
        f"{rates['encode']['gbs']:.0f} GB/s, decode {rates['decode']['gbs']:.0f} GB/s "
        f"({time.perf_counter() - t0:.1f} s)")
    return {"workload": f"single-root BFS + s4-bp128-d4 encode/decode of every level queue, 64 roots, "
                        f"Kronecker scale {scale}, 1x1 block",
            "hmean_gteps": hmean(teps) / 1e9, "mean_ms_per_root": float(np.mean(ms)),
            "codec": {"queues": len(queues), "values": sizes, "words": words,
                      "what": "level queues of 4 of the searches, each encoded and decoded once on the device",
                      **rates}}


def sub_c3(g5, g, roots):
    """C3 beside the headline: one rank's share of a 4 x 4 grid at scale 26
    (its 2^24 columns).  Per level of the seeded search of the first key, 4
    column-frontier bitmaps of that level's density within the rank's
    columns (262,144 words each) are OR-combined on the device, and the
    level's queue within the rank's columns goes through the codec; then 4
    predecessor arrays of 2^24 are min-reduced.  Rates from the device's own
    event timing of g5_vreduce_or / g5_vreduce_min_pred / g5_s4_encode /
    g5_s4_decode (OR (k+1)*8*W, min-pred (k+1)*8*n)."""
    t0 = time.perf_counter()
    n, k = 1 << 24, 4
    w = n // 64
    b = g5.BFS(g)
    b.run(int(roots[0]))
    pred, lev = b.result()
    b.free()
    rng = np.random.default_rng(26)
    maps = g5.DeviceArray(k * w, np.uint64)
    out = g5.DeviceArray(w, np.uint64)
    dens = []
    # code objects loaded before timing (the OR's and the codec's)
    g5.check(g5.lib().g5_vreduce_or(maps.ptr, k, w, out.ptr))
    for warm in (np.arange(0, 1 << 20, 3, dtype=np.uint32), np.arange(0, 3000, 3, dtype=np.uint32)):
        g5.s4_decode(g5.s4_encode(warm), warm.size)  # the long and the short streams' kernels
    g5.exchange_profile(True)
    for d in range(1, int(lev.max()) + 1):
        p = float(np.mean(lev[:n] == d))
        dens.append(p)
        bits = rng.random(k * n, dtype=np.float32) < p
        maps_h = np.packbits(bits, bitorder="little").view(np.uint64)
        g5.check(g5.lib().g5_memcpy_h2d(maps.ptr, g5._p(maps_h), maps_h.nbytes))
        g5.check(g5.lib().g5_vreduce_or(maps.ptr, k, w, out.ptr))
        q = np.flatnonzero(lev[:n] == d).astype(np.uint32)
        if q.size:
            back = g5.s4_decode(g5.s4_encode(q), q.size)
            assert np.array_equal(back, q), "codec round trip lost values"
    want = np.bitwise_or.reduce(maps_h.reshape(k, w), axis=0)
    assert np.array_equal(out.to_host(), want), "vreduce OR differs from numpy"
    # the predecessor combine: each of the 4 contributions holds the parents
    # of a random quarter of the reached columns (-1 elsewhere)
    own = rng.integers(0, k, size=n)
    preds_h = np.full((k, n), -1, np.int64)
    for j in range(k):
        sel = (own == j) & (pred[:n] >= 0)
        preds_h[j, sel] = pred[:n][sel]
    dp = g5.DeviceArray.from_host(preds_h.reshape(-1))
    dmin = g5.DeviceArray(n, np.int64)
    for _ in range(3):
        g5.check(g5.lib().g5_vreduce_min_pred(dp.ptr, k, n, dmin.ptr))
    assert np.array_equal(dmin.to_host(), pred[:n]), "min-pred combine differs from the search's parents"
    rates = _rates(g5.exchange_stats())
    g5.exchange_profile(False)
    for a in (maps, out, dp, dmin):
        a.free()
    log(f"c3: OR {rates['or']['gbs']:.0f} GB/s over {rates['or']['calls']} levels, min-pred "
        f"{rates['min_pred']['gbs']:.0f} GB/s, codec encode {rates['encode']['gbs']:.0f} / decode "
        f"{rates['decode']['gbs']:.0f} GB/s ({time.perf_counter() - t0:.1f} s)")
    return {"workload": "one rank's share of a 4x4 grid at Kronecker scale 26: 4 x 262,144-word column "
                        "bitmaps OR-combined per level (densities of the seeded search's levels within the "
                        "rank's 2^24 columns), the level queue through s4-bp128-d4, 4 x int64[2^24] "
                        "predecessors min-reduced",
            "level_densities": dens, **rates}


def bench_c2(args, g5):
    g = g5.Graph.kronecker(args.scale)
    roots = g.select_roots(1 << args.scale, args.roots)
    log(f"graph: scale {args.scale}, {g.nnz} nonzeros; roots {len(roots)}")
    b = g5.BFS(g)
    plain = []
    for r in roots.tolist():
        b.run(r)
        plain.append(b.nedge())
    for r in roots[:max(args.warmup, 1)].tolist():
        b.run(r, codec=True)
    b.profile(True)
    step_ms, per_root_ms, nedge = [], None, None
    for i in range(args.warmup + args.steps):
        g5.lib().g5_device_sync()
        t = time.perf_counter()
        ms_r, ne_r = [], []
        for r in roots.tolist():
            nl, ms = b.run(r, codec=True)
            ms_r.append(ms)
            ne_r.append(b.nedge())
        g5.lib().g5_device_sync()
        if i >= args.warmup:
            step_ms.append((time.perf_counter() - t) * 1e3)
            per_root_ms = ms_r if per_root_ms is None else [a + c for a, c in zip(per_root_ms, ms_r)]
        nedge = np.array(ne_r, np.float64)
    assert np.array_equal(nedge, np.array(plain, np.float64)), "codec in the loop changed a search"
    per_root_ms = np.array(per_root_ms) / args.steps
    teps = nedge / (per_root_ms / 1e3)
    value = hmean(teps) / 1e9
    kst1 = b.kernel_stats()
    kname = max(kst1, key=lambda k: kst1[k]["ms"])
    st = kst1[kname]
    gbs = st["bytes"] / (st["ms"] / 1e3) / 1e9 if st["ms"] > 0 else 0.0
    log(f"c2: hmean {value:.3f} GTEPS with the codec on every level, mean {per_root_ms.mean():.3f} ms/root")
    b.free()
    out = {
        "metric": METRIC, "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": float(np.mean(step_ms)), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 ids / u64 bitmaps (integer)",
        "data": "synthetic: Graph500 Kronecker (A=.57 B=C=.19), edgefactor 16, seeds (1,2), generated on the device",
        "config": {"workload": f"single-root BFS + s4-bp128-d4 encode/decode of every level queue, 64 roots, "
                               f"Kronecker scale {args.scale}, 1x1 block",
                   "scale": args.scale, "edgefactor": 16, "roots": int(len(roots)), "nnz": int(g.nnz)},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None, "kernel": kname,
                     "avg_launch_ms": st["ms"] / max(st["launches"], 1),
                     "algorithmic_bytes_per_launch": st["bytes"] / max(st["launches"], 1)},
    }
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline_codec(g, args.scale, roots.tolist(), nedge)
    return out


def bench_c1(args, g5):
    g = g5.Graph.kronecker(args.scale)
    roots = g.select_roots(1 << args.scale, args.roots)
    log(f"graph: scale {args.scale}, {g.nnz} nonzeros; roots {len(roots)}")
    b = g5.BFS(g)
    for r in roots[:max(args.warmup, 1)].tolist():
        b.run(r)
    b.profile(True)
    step_ms, per_root_ms, nedge = [], None, None
    for i in range(args.warmup + args.steps):
        g5.lib().g5_device_sync()
        t = time.perf_counter()
        ms_r, ne_r = [], []
        for r in roots.tolist():
            nl, ms = b.run(r)
            ms_r.append(ms)
            ne_r.append(b.nedge())
        g5.lib().g5_device_sync()
        if i >= args.warmup:
            step_ms.append((time.perf_counter() - t) * 1e3)
            per_root_ms = ms_r if per_root_ms is None else [a + c for a, c in zip(per_root_ms, ms_r)]
        nedge = np.array(ne_r, np.float64)
    per_root_ms = np.array(per_root_ms) / args.steps
    value = hmean(nedge / (per_root_ms / 1e3)) / 1e9
    kst1 = b.kernel_stats()
    kname = max(kst1, key=lambda k: kst1[k]["ms"])
    st = kst1[kname]
    gbs = st["bytes"] / (st["ms"] / 1e3) / 1e9 if st["ms"] > 0 else 0.0
    b.free()
    m = g5.MSBFS(g)
    for _ in range(max(args.warmup, 1)):
        m.run(roots)
    bms = [m.run(roots)[1] for _ in range(max(args.steps, 1))]
    ne_b = m.nedge().astype(np.float64)
    assert np.array_equal(ne_b, nedge), "batched and single-root TEPS numerators differ"
    t_batch = float(np.mean(bms))
    batched = hmean(ne_b / (t_batch / 1e3 / len(roots))) / 1e9
    log(f"c1: single-root hmean {value:.3f} GTEPS ({per_root_ms.mean():.3f} ms/root), batched {t_batch:.3f} ms "
        f"per {len(roots)}-root search, {batched:.2f} GTEPS")
    out = {
        "metric": METRIC, "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": float(np.mean(step_ms)), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 ids / u64 bitmaps (integer)",
        "data": "synthetic: Graph500 Kronecker (A=.57 B=C=.19), edgefactor 16, seeds (1,2), generated on the device",
        "config": {"workload": f"single-root BFS of 64 roots, Kronecker scale {args.scale}, 1x1 block "
                               f"(the CPU-runnable case)",
                   "scale": args.scale, "edgefactor": 16, "roots": int(len(roots)), "nnz": int(g.nnz)},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None, "kernel": kname,
                     "avg_launch_ms": st["ms"] / max(st["launches"], 1),
                     "algorithmic_bytes_per_launch": st["bytes"] / max(st["launches"], 1)},
        "batched_ms": t_batch, "batched_hmean_gteps": batched,
    }
    if args.cpu_baseline == "on":
        # every root, whatever the budget: the case is small enough
        out["cpu_baseline"] = cpu_baseline(g, args.scale, roots, nedge, budget_s=float("inf"))
        log(f"c1: CPU reference algorithm {out['cpu_baseline']['value']:.4f} GTEPS over all roots")
    return out


def bench_c3(args):
    """The driver runs as a child process; this process touches the GPU only
    afterwards (for the CPU baseline's graph)."""
    from gpugraph500_amd import build as gbuild
    cmd = [gbuild.DRIVER, "-s", str(args.scale), "-R", "4", "-C", "4", "-n", str(args.roots), "-x", "device",
           "-be", "s4-bp128-d4", "-share", "0", "0"]
    log("c3: " + " ".join(cmd))
    t0 = time.perf_counter()
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"g500_hip exited {r.returncode}: {r.stderr[-500:]}")
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    wall = time.perf_counter() - t0
    share_s = rec["share_expand_s"] + rec["share_column_s"] + rec["share_fold_s"]
    kx_gbs = rec["share_kexpand_bytes"] / (rec["share_kexpand_ms"] / 1e3) / 1e9 if rec["share_kexpand_ms"] else 0.0
    log(f"c3: emulated grid {rec['hmean_teps'] / 1e9:.3f} GTEPS; rank (0,0) share {share_s * 1e3:.2f} ms per search")
    out = {
        "metric": METRIC, "value": rec["hmean_teps"] / 1e9, "unit": "GTEPS", "n_gpus": 1, "steps": rec["roots"],
        "warmup": 1, "ms_per_step": rec["mean_time_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 ids / u64 bitmaps (integer)",
        "data": "synthetic: Graph500 Kronecker (A=.57 B=C=.19), edgefactor 16, seeds (1,2), generated on the device",
        "config": {"workload": f"4x4 grid emulated on one GPU (16 blocks of 2^24 x 2^24 at scale {args.scale}), "
                               f"device exchange, s4-bp128-d4 on the fold, rank (0,0)'s share timed",
                   "scale": args.scale, "grid": [4, 4], "roots": rec["roots"]},
        "rank_share_ms": {"expand": rec["share_expand_s"] * 1e3, "column_or": rec["share_column_s"] * 1e3,
                          "fold": rec["share_fold_s"] * 1e3, "total": share_s * 1e3},
        "exchange_words": {"raw": rec["raw_words"], "coded": rec["coded_words"]},
        "roofline": {"bound": "hbm", "achieved": kx_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kx_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "k_expand (rank 0,0)",
                     "avg_launch_ms": rec["share_kexpand_ms"] / max(rec["share_kexpand_launches"], 1),
                     "algorithmic_bytes_per_launch": rec["share_kexpand_bytes"] / max(rec["share_kexpand_launches"], 1)},
        "driver_wall_s": wall,
    }
    if args.cpu_baseline == "on":
        import gpugraph500_amd as g5
        g5.init(0)
        g = g5.Graph.kronecker(args.scale)
        roots = g.select_roots(1 << args.scale, 2)
        b = g5.BFS(g)
        nedge = []
        for r in roots.tolist():
            b.run(int(r))
            nedge.append(b.nedge())
        out["cpu_baseline"] = cpu_baseline(g, args.scale, roots, nedge)
        b.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", choices=["c4", "c2", "c3", "c1"], default="c4")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--traffic", choices=["auto", "off"], default="auto",
                    help="c4: measure roofline.traffic with rocprofv3 --pmc child passes")
    ap.add_argument("--levels", action="store_true", help="print per-level kernel accounting to stderr")
    ap.add_argument("--validate", choices=["on", "off"], default="on",
                    help="validate every measured search on the device (c4)")
    ap.add_argument("--place", type=int, default=4,
                    help="c4: fresh allocations of the parent records tried before timing (0: none)")
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="c4: pick the batched search's shape parameters before timing")
    ap.add_argument("--subconfigs", choices=["on", "off"], default="on",
                    help="c4: also measure the C2 (codec) and C3 (vreduce) sub-records")
    ap.add_argument("--pmc-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tune-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tune-json", default="", help=argparse.SUPPRESS)
    ap.add_argument("--probe", choices=["single", "exchange", "batched"], default=None,
                    help="a workload for a profiler run (see probe())")
    args = ap.parse_args()
    if args.scale is None:
        args.scale = DEFAULT_SCALE[args.config]
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the 2D row/column exchange "
                                   "is emulated on a single device",
                          "n_gpus": args.gpus}))
        return 2
    if args.pmc_probe:
        return pmc_probe(args)
    if args.tune_probe:
        return tune_probe(args)
    if args.probe:
        return probe(args)
    if args.config == "c3":
        print(json.dumps(bench_c3(args)))
        return 0
    traffic, traffic_info, probe_tuned = None, None, None
    if args.config == "c4" and args.traffic == "auto":
        # child processes, before this process initialises the GPU
        traffic, traffic_info, probe_tuned = measure_traffic(args)

    import gpugraph500_amd as g5

    g5.init(0)
    if args.config == "c4":
        out = bench_c4(args, g5, traffic, traffic_info, probe_tuned)
    elif args.config == "c1":
        out = bench_c1(args, g5)
    else:
        out = bench_c2(args, g5)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
