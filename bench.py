#!/usr/bin/env python3
"""Headline benchmark: CSR+CSC build Medges/s on one MI355X, with BFS GTEPS and
PageRank GTEPS (FP32 compensated vs FP64) on the same graph.

Workload (default): R-MAT scale 24, edge factor 16 (16,777,216 vertices,
268,435,456 uint32 <src,dst> records), a/b/c/d = .57/.19/.19/.05, seed 2,
split into P = 8 CPBT partitions (eta = 1, m = 1, S = 1) on one GPU.  A step
is one full build from the device-resident edge list: degree histogram, CPBT
boundaries, per-rank shuffle sort, two-level CSC and CSR with their indexes
(the reference's timed prep region).  Inputs are generated on the GPU before
the timed region.

Beside the headline line's value: BFS, PageRank and CC on the same graph;
the batched programs (R-MAT 22, 64 roots; 256 ER graphs) and the PPR
adjoint; R-MAT 16 PageRank; and the twitter-shaped graph (41,652,230
vertices, 1,468,365,182 edges): its build, BFS and FP32C-vs-FP64 PageRank
with the FP32C ranks' relative L1 against the FP64 oracle.  The host-CPU
baseline is the oracle (oracle/cpu_ref.cpp, OpenMP) on every host core this
process may use: the full R-MAT 24 build (the line's cpu_baseline), its BFS,
PageRank and CC, R-MAT 16 PageRank and the twitter-shaped build.

Prints one JSON line (rank 0).  Multi-GPU runs are not part of this build.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

WORKLOADS = {
    # name: (description, V, E, generator, P, S, eta, m)
    "rmat24": ("R-MAT scale 24, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 2; P=8 CPBT partitions, S=1",
               1 << 24, 1 << 28, ("rmat", 24, 2), 8, 1, 1, 1),
    "twitter": ("twitter-shaped power law (out alpha 1.9, in alpha 2.1), seed 3; P=1, S=1",
                41_652_230, 1_468_365_182, ("powerlaw", 3), 1, 1, 1, 20),
    "rmat16": ("R-MAT scale 16, edge factor 16, seed 1; P=1, S=1",
               1 << 16, 1 << 20, ("rmat", 16, 1), 1, 1, 1, 20),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def median_time(fn, runs=3, warm=1):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def file_load_extras(eg, edges, E, V, P, S, eta, m):
    """The step before the build: the binary edge list (the reference's file
    format) read into HBM by the pipelined reader, page cache warm; then the
    same file loaded and built in one call, each rank slice's first sorts
    running while the next slice is read, against the read followed by the
    build."""
    import tempfile
    import ctypes as C
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    try:
        edges.download(eg.REC8, E).tofile(path)
        dst = eg.DeviceBuffer(E * 8)
        n = C.c_uint64()

        def load():
            eg.check(eg.lib().eg_load_edge_file(path.encode(), 8, C.c_void_p(dst.ptr), E, C.byref(n), None))

        t = median_time(load)
        g = eg.Graph(V, P, S, eta, m, 8)

        def load_then_build():
            load()
            g.build(dst.ptr, E)

        t_seq = median_time(load_then_build)
        t_ovl = median_time(lambda: g.build_file(path, dst.ptr, E))
        del g
    finally:
        os.unlink(path)
    log(f"[bench] edge file load: {E * 8 / t / 1e9:.2f} GB/s ({t * 1e3:.1f} ms for {E * 8 / 2**30:.1f} GiB); "
        f"load + build {t_seq * 1e3:.1f} ms, overlapped {t_ovl * 1e3:.1f} ms")
    return {"bytes": E * 8, "ms": round(t * 1e3, 2), "gbs": round(E * 8 / t / 1e9, 2),
            "load_then_build_ms": round(t_seq * 1e3, 2), "load_build_overlapped_ms": round(t_ovl * 1e3, 2),
            "note": "4 reader threads, 32 MiB pread chunks into per-thread pinned buffers kept between loads, "
                    "H2D on per-thread streams overlapped with the reads; file in the page cache; overlapped = "
                    "eg_graph_build_file (each rank slice's symmetrisation and first sorts queued as it lands)"}


def ppr_extras(eg, iterations=250, d=0.85):
    """Personalised PageRank forward and adjoint solves on R-MAT scale 22 with
    fp32 weights U[0.1,1] row-normalised (FP64 weights per edge slot), p
    Dirichlet(1) on 1024 seeds, dL/dx ~ N(0,1); `iterations` fixed-point steps
    each (the gradient check against central differences is a test)."""
    scale, V = 22, 1 << 22
    E = 16 * V
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=22, rec_bytes=12, w_lo=0.1, w_hi=1.0)
    g = eg.Graph(V, 1, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    s_src, _, s_w = g.ppr_slots()
    src = s_src.download(np.uint32, E).astype(np.int64)
    w = s_w.download(np.float32, E).astype(np.float64)
    w /= np.bincount(src, weights=w, minlength=V)[src]
    rng = np.random.default_rng(22)
    p = np.zeros(V)
    p[rng.choice(V, 1024, replace=False)] = rng.dirichlet(np.ones(1024))
    c = rng.standard_normal(V)
    dw, dp, dc = eg.DeviceBuffer.from_array(w), eg.DeviceBuffer.from_array(p), eg.DeviceBuffer.from_array(c)
    dx, gp, gw = eg.DeviceBuffer(V * 8), eg.DeviceBuffer(V * 8), eg.DeviceBuffer(E * 8)
    tf = median_time(lambda: g.ppr_forward(dw.ptr, dp.ptr, d, iterations, dx.ptr))
    ta = median_time(lambda: g.ppr_adjoint(dw.ptr, dx.ptr, dc.ptr, d, iterations, gp.ptr, gw.ptr))
    log(f"[bench] ppr (R-MAT 22, {iterations} steps): forward {tf * 1e3:.1f} ms, adjoint {ta * 1e3:.1f} ms")
    return {"workload": "R-MAT scale 22, weights U[0.1,1] row-normalised, p Dirichlet(1) on 1024 seeds",
            "iterations": iterations, "ms_forward": round(tf * 1e3, 2), "ms_adjoint": round(ta * 1e3, 2),
            "gteps_forward": round(iterations * E / tf / 1e9, 2), "gteps_adjoint": round(iterations * E / ta / 1e9, 2)}


def batching_extras(eg):
    """Batched programs: 64 seeded roots of BFS and SSSP on R-MAT scale 22 with
    fp32 weights U[0,1) (one batched run against 64 single-root runs), and the
    CSR build of 256 Erdos-Renyi graphs (|V| log-uniform in 1e3..1e6, average
    degree U[8,32]) as one batch against one build per graph."""
    out = {}
    scale, V = 22, 1 << 22
    E = 16 * V
    buf = eg.DeviceBuffer(E * 12)
    eg.gen_rmat(buf.ptr, E, scale, seed=4, rec_bytes=12, w_lo=0.0, w_hi=1.0)
    g = eg.Graph(V, 1, 1, 1, 1, 12)
    g.build(buf.ptr, E)
    g.prepare_engine()
    tune, tms = g.autotune_engine(n_roots=64, seed=4)  # untimed; its own sample roots
    out["engine_tune"] = {"picked": tune, "best_ms": [round(x, 3) for x in tms]}
    log(f"[bench] engine tune (R-MAT 22 weighted): {tune}")
    roots = np.random.default_rng(4).integers(0, V, 64).astype(np.uint32)
    parent = eg.DeviceBuffer(V * 64 * 4)
    depth = eg.DeviceBuffer(V * 64 * 4)
    res = {}

    def bfs_batch():
        res["b"] = g.bfs_multi(roots, eg.POLICY_AUTO, parent, depth, download=False)

    def bfs_single():
        p1, d1 = eg.DeviceBuffer(V * 4), eg.DeviceBuffer(V * 4)
        for r in roots:
            g.bfs(int(r), eg.POLICY_AUTO, p1, d1, download=False)

    tb = median_time(bfs_batch)
    ts = median_time(bfs_single, runs=1)
    levels, edges, steps = res["b"]
    out["bfs_64_roots"] = {"workload": "R-MAT scale 22, edge factor 16, fp32 weights U[0,1), seed 4; 64 roots, seed 4",
                           "ms_batched": round(tb * 1e3, 3), "ms_64_single": round(ts * 1e3, 3),
                           "speedup": round(ts / tb, 2), "steps": int(steps),
                           "gteps_batched": round(float(edges.sum()) / tb / 1e9, 3)}
    log(f"[bench] bfs 64 roots: batched {tb * 1e3:.2f} ms vs single {ts * 1e3:.2f} ms")
    del parent, depth
    dist = eg.DeviceBuffer(V * 64 * 4)
    tb = median_time(lambda: g.sssp_multi(roots, eg.POLICY_AUTO, dist, download=False))

    def sssp_single():
        d1 = eg.DeviceBuffer(V * 4)
        for r in roots:
            g.sssp(int(r), eg.POLICY_AUTO, d1, download=False)

    ts = median_time(sssp_single, runs=1)
    out["sssp_64_roots"] = {"ms_batched": round(tb * 1e3, 3), "ms_64_single": round(ts * 1e3, 3),
                            "speedup": round(ts / tb, 2)}
    log(f"[bench] sssp 64 roots: batched {tb * 1e3:.2f} ms vs single {ts * 1e3:.2f} ms")
    del dist, g, buf
    out["ppr_adjoint"] = ppr_extras(eg)
    rng = np.random.default_rng(256)
    n = 256
    nv = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n)).astype(np.uint32)
    ne = (nv.astype(np.float64) * rng.uniform(8, 32, n)).astype(np.uint64)
    total = int(ne.sum())
    ebuf = eg.DeviceBuffer(total * 8)
    eg.gen_er_batch(ebuf.ptr, nv, ne, seed=5)
    b = eg.BatchCSR(ebuf.ptr, nv, ne)
    tb = median_time(lambda: b.rebuild(ebuf.ptr))
    off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
    singles = [eg.BatchCSR(ebuf.ptr + int(off[i]) * 8, nv[i:i + 1], ne[i:i + 1]) for i in range(n)]

    def one_by_one():
        for i in range(n):
            singles[i].rebuild(ebuf.ptr + int(off[i]) * 8)

    ts = median_time(one_by_one, runs=1)
    out["er_256_csr"] = {"workload": "256 Erdos-Renyi graphs, |V| log-uniform 1e3..1e6, degree U[8,32], seed 256",
                         "vertices": int(nv.sum()), "edges": total, "ms_batched": round(tb * 1e3, 3),
                         "ms_one_by_one": round(ts * 1e3, 3), "speedup": round(ts / tb, 2),
                         "medges_per_s_batched": round(total / tb / 1e6, 1)}
    log(f"[bench] 256 ER CSR builds: batched {tb * 1e3:.2f} ms vs one by one {ts * 1e3:.2f} ms")
    return out


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t0


def build_algorithmic_bytes(g, E, rec=8):
    """What a build must move at least: the COO list read once, the CSC and
    CSR lists written once, out- and in-degree, every sub-partition's dense
    CSR index (u64[V+1], the reference's layout) and bitmap, and the
    compressed CSC index entries (12 B each, sentinels included)."""
    V, Q = g.V, g.P * g.S
    cidx = sum(g.csc_ptrs(p, s)[3] for p in range(g.P) for s in range(g.S))
    return E * rec * 3 + V * 8 + Q * (V + 1) * 8 + Q * ((V >> 6) + 1) * 8 + cidx * 12


def pagerank_block(eg, g, iters, E, oracle_pr=None):
    """Plain FP64 (built CSC) against compensated FP32 on the degree-ordered
    view, 20 iterations each, median of 3 after a warm-up.  The view's default
    form is the relabelled CSC with P = S = 1 and rows sorted by new source id
    otherwise; with P = S = 1 the sorted-row form (costlier to derive, faster
    per pull) is timed too.  The view-counted ratio of the default form is
    the plain FP64 run against an FP32C run that starts without the view and
    derives it (the relabelled CSC's sources are then written by its first
    pull); the sorted-row form adds its one-time derivation to one run."""
    V = g.V
    out64, out32, scratch = eg.DeviceBuffer(V * 8), eg.DeviceBuffer(V * 8), eg.DeviceBuffer(V * 8)
    t64 = median_time(lambda: g.pagerank(iters, eg.PR_FP64, out64))
    firsts = []
    for _ in range(3):
        g.release_pagerank_view()  # untimed: frees the view
        t0 = time.perf_counter()
        g.pagerank(iters, eg.PR_FP32C, out32)  # synchronises
        firsts.append(time.perf_counter() - t0)
    t32_first = statistics.median(firsts)
    g.release_pagerank_view()
    view_ms = g.prepare_pagerank_view()  # the derivation on its own, for the record
    t32 = median_time(lambda: g.pagerank(iters, eg.PR_FP32C, out32))
    t64o = median_time(lambda: g.pagerank(iters, eg.PR_FP64 | eg.PR_ORDERED, scratch))
    t32p = median_time(lambda: g.pagerank(iters, eg.PR_FP32C | eg.PR_PLAIN, scratch))
    x64 = out64.download(np.float64, V)
    x32 = out32.download(np.float64, V)
    gt = lambda t: round(iters * E / t / 1e9, 3)
    one = g.P * g.S == 1
    out = {"iterations": iters, "view_form": "relabelled CSC" if one else "sorted rows",
           "view_prepare_ms": round(view_ms, 2),
           "ms_fp64": round(t64 * 1e3, 2), "ms_fp32c": round(t32 * 1e3, 2),
           "gteps_fp64": gt(t64), "gteps_fp32c": gt(t32),
           "speedup_fp32c": round(t64 / t32, 3),
           "ms_fp32c_first_run": round(t32_first * 1e3, 2),
           "speedup_fp32c_view_counted": round(t64 / t32_first, 3),
           "l1_fp32c_vs_fp64_gpu": float(np.abs(x32 - x64).sum() / np.abs(x64).sum()),
           "gteps_fp64_ordered": gt(t64o), "gteps_fp32c_plain": gt(t32p),
           "layouts": "fp64 = plain FP64 on the built CSC; fp32c = compensated FP32 on the degree-ordered view "
                      "(view_form); ms_fp32c_first_run = a 20-iteration FP32C run that starts without the view "
                      "and derives it; view_counted = ms_fp64 / ms_fp32c_first_run"}
    if one:
        sorted_ms = g.prepare_pagerank_view(eg.PR_SORTED)
        t32s = median_time(lambda: g.pagerank(iters, eg.PR_FP32C | eg.PR_SORTED, scratch))
        xs = scratch.download(np.float64, V)
        out["sorted_rows"] = {"view_prepare_ms": round(sorted_ms, 2), "ms_fp32c": round(t32s * 1e3, 2),
                              "gteps_fp32c": gt(t32s), "speedup_fp32c": round(t64 / t32s, 3),
                              "speedup_fp32c_view_counted": round(t64 / (t32s + sorted_ms * 1e-3), 3),
                              "l1_fp32c_vs_fp64_gpu": float(np.abs(xs - x64).sum() / np.abs(x64).sum())}
        g.prepare_pagerank_view()  # the default form again
    if oracle_pr is not None:
        n1 = np.abs(oracle_pr).sum()
        out["l1_fp32c_vs_fp64_oracle"] = float(np.abs(x32 - oracle_pr).sum() / n1)
        out["l1_fp64_vs_fp64_oracle"] = float(np.abs(x64 - oracle_pr).sum() / n1)
        if one:
            out["sorted_rows"]["l1_fp32c_vs_fp64_oracle"] = float(np.abs(xs - oracle_pr).sum() / n1)
    log(f"[bench] pagerank: fp64 {gt(t64)} GTEPS, fp32c {gt(t32)} GTEPS ({t64 / t32:.2f}x; "
        f"{out['speedup_fp32c_view_counted']:.2f}x for a run that derives the view, {t32_first * 1e3:.1f} ms; "
        f"the view alone {view_ms:.1f} ms)")
    if one:
        so = out["sorted_rows"]
        log(f"[bench] pagerank, sorted-row view: fp32c {so['gteps_fp32c']} GTEPS ({so['speedup_fp32c']:.2f}x, "
            f"{so['speedup_fp32c_view_counted']:.2f}x with the {so['view_prepare_ms']:.1f} ms view counted)")
    return out, x32


def bfs_block(eg, g, root):
    V = g.V
    parent, depth = eg.DeviceBuffer(V * 4), eg.DeviceBuffer(V * 4)
    res = {}

    def run():
        res["b"] = g.bfs(root, eg.POLICY_AUTO, parent, depth, download=False)

    tb = median_time(run)
    levels, trav = res["b"]
    log(f"[bench] bfs from {root}: {tb * 1e3:.2f} ms, {trav / tb / 1e9:.2f} GTEPS")
    return {"root": root, "policy": "auto", "levels": levels, "edges_traversed": trav,
            "ms": round(tb * 1e3, 3), "gteps": round(trav / tb / 1e9, 3)}


def cpu_block(name, seconds, work, unit_div, unit, cores):
    return {"value": round(work / seconds / unit_div, 3), "unit": unit, "s": round(seconds, 3), "cores": cores,
            "workload": name}


def rocpd_summary(root, pattern=None):
    """Markdown table of one profiled probe from rocprofv3's databases under
    root: root/kt (--kernel-trace --stats) for calls and average durations,
    and one directory per --pmc run for the counters, averaged per kernel
    over its dispatches.  FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are
    scaled to bytes by what they read for the probe's 1 GiB device copy
    (k_copy16: 1 GiB read and 1 GiB written)."""
    import glob
    import re
    import sqlite3

    def short(n):
        n = re.sub(r"\(anonymous namespace\)::", "", n)
        n = re.sub(r"^void ", "", n)
        return re.sub(r"^eg::", "", n).split("(")[0]

    stats, vgprs = {}, {}
    for db in glob.glob(os.path.join(root, "kt", "*.db"))[:1]:
        con = sqlite3.connect(db)
        for name, calls, tot, avg in con.execute("select name, total_calls, total_duration, average from top_kernels"):
            stats[short(name)] = (calls, tot / 1e3, avg / 1e3)  # the view's durations are in us
        for name, v, a in con.execute("select name, max(vgpr_count), max(accum_vgpr_count) from kernels group by name"):
            vgprs[short(name)] = (v or 0) + (a or 0)
    cnt = {}
    for db in sorted(glob.glob(os.path.join(root, "*", "*.db"))):
        if os.sep + "kt" + os.sep in db:
            continue
        con = sqlite3.connect(db)
        q = ("select kernel_name, counter_name, avg(v) from (select dispatch_id, kernel_name, counter_name, "
             "sum(value) v from counters_collection group by dispatch_id, counter_name) "
             "group by kernel_name, counter_name")
        for kn, cn, v in con.execute(q):
            cnt.setdefault(short(kn), {})[cn] = v
    copy = cnt.get("k_copy16", {})
    fs = (1 << 30) / (copy["FETCH_SIZE"] * 1024) if copy.get("FETCH_SIZE") else None
    ws = (1 << 30) / (copy["WRITE_SIZE"] * 1024) if copy.get("WRITE_SIZE") else None
    lines = [f"calibration (k_copy16, 1 GiB read + 1 GiB written): FETCH_SIZE x {fs}, WRITE_SIZE x {ws}", "",
             "| kernel | calls | avg ms | VGPRs | FETCH GB/dispatch | WRITE GB/dispatch | L2 hit % | VALU / wave | "
             "VMEM rd / wave | wait / busy |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    fmt = lambda x, f="{:.3f}": f.format(x) if x is not None else "-"
    for k in sorted(set(stats) | set(cnt), key=lambda k: -stats.get(k, (0, 0, 0))[1]):
        if pattern and not re.search(pattern, k):
            continue
        c, st = cnt.get(k, {}), stats.get(k, (0, 0.0, 0.0))
        f, w, h, m, waves = (c.get("FETCH_SIZE"), c.get("WRITE_SIZE"), c.get("TCC_HIT_sum"), c.get("TCC_MISS_sum"),
                             c.get("SQ_WAVES"))
        lines.append(
            f"| `{k}` | {st[0]} | {st[2]:.3f} | {vgprs.get(k, '-')} | {fmt(f * 1024 * fs / 1e9 if f and fs else None)} | "
            f"{fmt(w * 1024 * ws / 1e9 if w and ws else None)} | "
            f"{fmt(100 * h / (h + m) if h is not None and m else None, '{:.1f}')} | "
            f"{fmt(c['SQ_INSTS_VALU'] / waves if waves and c.get('SQ_INSTS_VALU') else None, '{:.0f}')} | "
            f"{fmt(c['SQ_INSTS_VMEM_RD'] / waves if waves and c.get('SQ_INSTS_VMEM_RD') else None, '{:.0f}')} | "
            f"{fmt(c['SQ_WAIT_INST_ANY'] / c['SQ_BUSY_CYCLES'] if c.get('SQ_BUSY_CYCLES') and c.get('SQ_WAIT_INST_ANY') else None, '{:.2f}')} |")
    return "\n".join(lines)


def twitter_block(eg, args, cores):
    """The twitter-shaped graph (41,652,230 vertices, 1,468,365,182 edges,
    seed 3, P = 1): build Medges/s, BFS GTEPS, FP32C vs FP64 PageRank with
    the FP32C ranks' relative L1 against the FP64 oracle (computed on the host,
    untimed) and the oracle's build timed on every host core."""
    import oracle
    desc, V, E, gen, P, S, eta, m = WORKLOADS["twitter"]
    edges = eg.DeviceBuffer(E * 8)
    eg.gen_powerlaw(edges.ptr, E, V, seed=gen[1])
    g = eg.Graph(V, P, S, eta, m, 8)
    if not args.no_autotune:
        g.autotune_sort(edges.ptr, E)
    for _ in range(1):
        g.build(edges.ptr, E)
    eg.synchronize()
    t0 = time.perf_counter()
    steps = 3
    for _ in range(steps):
        g.build(edges.ptr, E)
    eg.synchronize()
    tb = (time.perf_counter() - t0) / steps
    bb = build_algorithmic_bytes(g, E)
    out = {"workload": desc, "vertices": V, "edges": E, "build_ms": round(tb * 1e3, 2),
           "build_medges_per_s": round(E / tb / 1e6, 1), "build_steps": steps,
           "sort_variant": dict(zip(("scheme", "threads", "digit_bits"), eg.get_sort_variant(8))),
           "build_level": {"algorithmic_bytes_per_build": int(bb), "achieved_gbs": round(bb / tb / 1e9, 1),
                           "frac": round(bb / tb / 1e9 / HBM_PEAK_GBS, 4)}}
    log(f"[bench] twitter build: {tb * 1e3:.1f} ms, {E / tb / 1e6:.0f} Medges/s, "
        f"{out['build_level']['frac']:.3f} of peak as a whole")
    prep_ms = g.prepare_engine()
    tv_ms = g.prepare_traversal_view()
    if not args.no_autotune:
        tune, tms = g.autotune_engine(n_roots=16, seed=3)  # untimed
        out["engine_tune"] = {"picked": tune, "best_ms": [round(x, 3) for x in tms]}
    deg = g.out_degree()
    out["bfs"] = bfs_block(eg, g, int(np.argmax(deg)))
    del deg
    # oracle on the host: the CPU build (timed) and the FP64 reference ranks
    host = edges.download(eg.REC8, E)
    del edges
    og = oracle.OracleGraph(host, V, P=P, S=S, eta=eta, m=m, threads=cores)
    del host
    ref, t_pr_cpu = timed(lambda: og.pagerank(20)[0])
    cpu_build = og.build_seconds
    del og
    pr, _ = pagerank_block(eg, g, 20, E, oracle_pr=ref)
    out["pagerank"] = pr
    out["engine_ready_ms"] = round(tb * 1e3 + prep_ms + tv_ms + pr["view_prepare_ms"], 2)
    out["engine_ready_parts_ms"] = {"build": round(tb * 1e3, 2), "engine_prepare": round(prep_ms, 2),
                                    "traversal_view": round(tv_ms, 2), "pagerank_view": pr["view_prepare_ms"]}
    out["cpu"] = {"build": cpu_block("oracle build (split_graph-equivalent region)", cpu_build, E, 1e6,
                                     "Medges/s", cores),
                  "pagerank_fp64": cpu_block("oracle FP64 PageRank, 20 iterations", t_pr_cpu, 20 * E, 1e9, "GTEPS",
                                             cores)}
    out["build_vs_cpu"] = round(out["build_medges_per_s"] / out["cpu"]["build"]["value"], 1)
    out["targets"] = {"build_vs_host_cpu_ge_20x": out["build_vs_cpu"] >= 20,
                      "fp32c_vs_fp64_ge_1.6x": pr["speedup_fp32c"] >= 1.6,
                      "fp32c_vs_fp64_ge_1.6x_view_counted": pr["speedup_fp32c_view_counted"] >= 1.6,
                      "fp32c_sorted_rows_vs_fp64_ge_1.6x": pr["sorted_rows"]["speedup_fp32c"] >= 1.6,
                      "fp32c_l1_vs_fp64_oracle_le_1e-6": max(pr["l1_fp32c_vs_fp64_oracle"],
                                                             pr["sorted_rows"]["l1_fp32c_vs_fp64_oracle"]) <= 1e-6}
    log(f"[bench] twitter: build {out['build_vs_cpu']}x the host CPU ({cores} cores); "
        f"FP32C L1 vs FP64 oracle {pr['l1_fp32c_vs_fp64_oracle']:.2e}")
    del g
    return out


def rmat16_block(eg, cores):
    """R-MAT scale 16 (seed 1, P = 1): 20 PageRank iterations in FP64 on the
    GPU and on the host (the reference's plain CPU configuration)."""
    import oracle
    desc, V, E, gen, P, S, eta, m = WORKLOADS["rmat16"]
    edges = eg.DeviceBuffer(E * 8)
    eg.gen_rmat(edges.ptr, E, gen[1], seed=gen[2])
    g = eg.Graph(V, P, S, eta, m, 8)
    tb = median_time(lambda: g.build(edges.ptr, E))
    og = oracle.OracleGraph(edges.download(eg.REC8, E), V, P=P, S=S, eta=eta, m=m, threads=cores)
    ref, t_cpu = timed(lambda: og.pagerank(20)[0])
    out64 = eg.DeviceBuffer(V * 8)
    t64 = median_time(lambda: g.pagerank(20, eg.PR_FP64, out64))
    x = out64.download(np.float64, V)
    return {"workload": desc, "build_ms": round(tb * 1e3, 3), "cpu_build_s": round(og.build_seconds, 4),
            "pagerank_fp64_ms": round(t64 * 1e3, 3), "pagerank_fp64_gteps": round(20 * E / t64 / 1e9, 3),
            "cpu_pagerank_fp64": cpu_block("oracle FP64 PageRank, 20 iterations", t_cpu, 20 * E, 1e9, "GTEPS", cores),
            "l1_fp64_vs_oracle": float(np.abs(x - ref).sum() / np.abs(ref).sum())}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--no-engine", action="store_true", help="skip the BFS / PageRank / CC figures")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batching", action="store_true", help="skip the batched BFS/SSSP/CSR figures")
    ap.add_argument("--no-twitter", action="store_true", help="skip the twitter-shaped graph's block")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: two builds of the workload, then a 1 GiB device copy (the "
                         "calibration of FETCH_SIZE / WRITE_SIZE); prints no bench line")
    ap.add_argument("--rocpd-summary", metavar="DIR",
                    help="print the markdown summary of a profiled probe's rocprofv3 databases under DIR "
                         "(DIR/kt: kernel trace; DIR/<name>: one --pmc run each) and exit")
    ap.add_argument("--kernels", metavar="REGEX", help="with --rocpd-summary: only kernels matching REGEX")
    ap.add_argument("--engine-probe", action="store_true",
                    help="for rocprofv3 runs: build the workload and its views, then two runs each of plain FP64 "
                         "and ordered FP32C PageRank (20 iterations) and of BFS from the max-out-degree vertex, "
                         "then a 1 GiB device copy (the counter calibration); prints no bench line")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default sort variant")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every host core this process may use")
    args = ap.parse_args()
    if args.rocpd_summary:
        print(rocpd_summary(args.rocpd_summary, args.kernels))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one MI355X only; multi-GPU runs (--gpus > 1) are not part of it",
              file=sys.stderr)
        sys.exit(2)

    import endgraph_amd as eg

    cores = args.cpu_threads or host_cores()
    desc, V, E, gen, P, S, eta, m = WORKLOADS[args.workload]
    log(f"[bench] workload {args.workload}: V={V} E={E} P={P} S={S}")
    edges = eg.DeviceBuffer(E * 8)
    if gen[0] == "rmat":
        eg.gen_rmat(edges.ptr, E, gen[1], seed=gen[2])
    else:
        eg.gen_powerlaw(edges.ptr, E, V, seed=gen[1])
    eg.synchronize()
    log("[bench] edges generated")

    g = eg.Graph(V, P, S, eta, m, 8)
    if args.traffic_probe:
        for _ in range(2):
            g.build(edges.ptr, E)
        a, b = eg.DeviceBuffer(1 << 30), eg.DeviceBuffer(1 << 30)
        eg.memcpy_d2d(b.ptr, a.ptr, 1 << 30)
        eg.synchronize()
        log("[bench] traffic probe done: 2 builds, one 1 GiB device copy")
        return
    if args.engine_probe:
        g.build(edges.ptr, E)
        g.prepare_engine()
        g.prepare_traversal_view()
        g.prepare_pagerank_view()
        root = int(np.argmax(g.out_degree()))
        out = eg.DeviceBuffer(V * 8)
        pulls = {}
        for _ in range(2):
            for name, prec in (("fp64_plain", eg.PR_FP64), ("fp32c_ordered", eg.PR_FP32C)):
                eg.profile_reset()
                eg.profile_enable(True)
                g.pagerank(20, prec, out)
                eg.profile_enable(False)
                ms, n, b = eg.profile_read("pr_pull")
                pulls[name] = {"launches": n, "avg_ms": round(ms / n, 4), "algorithmic_bytes_per_launch": b // n,
                               "gbs": round(b / n / (ms / n * 1e-3) / 1e9, 1)}
            levels, trav = g.bfs(root, eg.POLICY_AUTO, download=False)
        del out
        eg.synchronize()
        print(json.dumps({"probe": "engine", "workload": desc, "vertices": V, "edges": E, "pulls": pulls,
                          "bfs": {"root": root, "levels": levels, "edges_traversed": trav}}), flush=True)
        a, b = eg.DeviceBuffer(1 << 30), eg.DeviceBuffer(1 << 30)
        eg.memcpy_d2d(b.ptr, a.ptr, 1 << 30)
        eg.synchronize()
        log("[bench] engine probe done: 2 x (PageRank FP64, PageRank FP32C, BFS), one 1 GiB device copy")
        return
    variant = None
    if not args.no_autotune:
        sc, nt, bits, ms = g.autotune_sort(edges.ptr, E)  # untimed: every variant builds twice
        variant = {"scheme": ["three-launch", "one-pass", "three-launch-xcd"][sc], "threads": nt,
                   "tile_records": nt * 16, "digit_bits": bits, "autotune_best_ms": round(ms, 2)}
        log(f"[bench] autotuned sort: {variant}")
    for i in range(args.warmup):
        g.build(edges.ptr, E)
        log(f"[bench] warmup build {i} done")
    eg.profile_reset()
    eg.profile_enable(True)
    eg.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        g.build(edges.ptr, E)  # synchronises at the end of every build
    eg.synchronize()
    t = time.perf_counter() - t0
    eg.profile_enable(False)
    ms_per_step = 1e3 * t / args.steps
    medges = E * args.steps / t / 1e6
    log(f"[bench] build: {ms_per_step:.2f} ms/step, {medges:.1f} Medges/s")

    kernels = {}
    for k in ("radix_downsweep", "radix_upsweep", "scan", "degree", "cindex", "csr_index"):
        ms, n, b = eg.profile_read(k)
        if n:
            kernels[k] = {"ms_total": ms, "launches": n, "avg_ms": ms / n, "bytes_per_launch": b / n,
                          "gbs": (b / n) / (ms / n * 1e-3) / 1e9}
    dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
    dk = kernels[dom]
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.workload}.json")
    if os.path.exists(tf):
        tj = json.load(open(tf))
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_src = (f"copied from {os.path.relpath(tf, ROOT)} (calibrated rocprofv3 --pmc FETCH_SIZE and "
                       f"WRITE_SIZE of {tj.get('source', 'a separate run')}), not measured in this run")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(dk["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dk["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "bytes_per_launch": int(dk["bytes_per_launch"]), "avg_launch_ms": round(dk["avg_ms"], 4),
                "launches": dk["launches"]}
    # what a plain streaming copy of the same bytes reaches on this device
    # (read + write, 16-byte vectors): the practical ceiling of a pass that
    # reads and writes every record once
    half = int(dk["bytes_per_launch"]) // 2
    ca, cb = eg.DeviceBuffer(half), eg.DeviceBuffer(half)
    eg.memcpy_d2d(cb.ptr, ca.ptr, half)
    eg.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        eg.memcpy_d2d(cb.ptr, ca.ptr, half)
    eg.synchronize()
    copy_gbs = 2 * half * 10 / (time.perf_counter() - t0) / 1e9
    del ca, cb
    roofline["copy_gbs"] = round(copy_gbs, 1)
    roofline["frac_of_copy"] = round(dk["gbs"] / copy_gbs, 4)
    bb = build_algorithmic_bytes(g, E)
    roofline["build_level"] = {"algorithmic_bytes_per_build": int(bb),
                               "achieved_gbs": round(bb / (ms_per_step * 1e-3) / 1e9, 1),
                               "frac": round(bb / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                               "bytes": "COO read + CSC and CSR written + degrees + dense CSR indexes and bitmaps + "
                                        "compressed CSC index"}
    log(f"[bench] {dom}: {dk['gbs']:.0f} GB/s = {dk['gbs'] / HBM_PEAK_GBS:.3f} of peak, "
        f"{dk['gbs'] / copy_gbs:.3f} of a streaming copy ({copy_gbs:.0f} GB/s); build as a whole "
        f"{roofline['build_level']['frac']:.3f} of peak")

    extra = {"kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in d.items()}
                         for k, d in kernels.items()}}
    cpu_extra = {"host": {"cores": cores, "model": host_cpu_model()}}
    og = None
    if not args.no_cpu_baseline:
        import oracle
        host = edges.download(eg.REC8, E)
        og = oracle.OracleGraph(host, V, P=P, S=S, eta=eta, m=m, threads=cores)
        cpu_extra["build"] = cpu_block("oracle build (split_graph-equivalent region), full list", og.build_seconds,
                                       E, 1e6, "Medges/s", cores)
        log(f"[bench] cpu build: {cpu_extra['build']['value']} Medges/s on {cores} threads")

    if not args.no_engine:
        prep_ms = g.prepare_engine()
        tv_ms = g.prepare_traversal_view()
        extra["engine_prepare_ms"] = round(prep_ms, 2)
        extra["traversal_view_prepare_ms"] = round(tv_ms, 2)
        if not args.no_autotune:
            tune, tms = g.autotune_engine(n_roots=16, seed=2)  # untimed
            extra["engine_tune"] = {"picked": tune, "best_ms": [round(x, 3) for x in tms]}
            log(f"[bench] engine tune: {tune}")
        root = int(np.argmax(g.out_degree()))
        extra["bfs"] = bfs_block(eg, g, root)
        ref = None
        if og is not None:
            (o_par, _, _), t_bfs = timed(lambda: og.bfs(root, 1))
            cpu_extra["bfs"] = cpu_block("oracle BFS (direction-optimising) from the same root", t_bfs,
                                         extra["bfs"]["edges_traversed"], 1e9, "GTEPS", cores)
            del o_par
            ref, t_pr = timed(lambda: og.pagerank(20)[0])
            cpu_extra["pagerank_fp64"] = cpu_block("oracle FP64 PageRank, 20 iterations", t_pr, 20 * E, 1e9,
                                                   "GTEPS", cores)
        extra["pagerank"], _ = pagerank_block(eg, g, 20, E, oracle_pr=ref)
        extra["engine_ready_ms"] = round(ms_per_step + prep_ms + tv_ms + extra["pagerank"]["view_prepare_ms"], 2)
    del og

    if not args.no_engine and args.workload == "rmat24":
        # CC on the same list symmetrised with reverse edges (2E records), as
        # load_undirected_from_directed builds it
        gs = eg.Graph(V, P, S, eta, m, 8)
        t_sym = median_time(lambda: gs.load_undirected_from_directed(edges.ptr, E), runs=3)
        gs.prepare_engine()
        label = eg.DeviceBuffer(V * 4)
        res = {}

        def run_cc():
            res["cc"] = gs.cc(eg.POLICY_REFERENCE, label, download=False)

        tc = median_time(run_cc)
        comps = int(np.unique(label.download(np.uint32, V)).size)
        extra["cc"] = {"graph": "symmetrised (2E records)", "build_ms": round(t_sym * 1e3, 2),
                       "build_medges_per_s": round(2 * E / t_sym / 1e6, 1), "ms": round(tc * 1e3, 3),
                       "iterations": res["cc"][1], "components": comps,
                       "gteps": round(res["cc"][1] * 2 * E / tc / 1e9, 2)}
        log(f"[bench] cc (symmetrised): build {t_sym * 1e3:.1f} ms, cc {tc * 1e3:.2f} ms, "
            f"{res['cc'][1]} iterations, {comps} components")
        del gs, label
        if not args.no_cpu_baseline:
            import oracle
            ogs = oracle.OracleGraph(edges.download(eg.REC8, E), V, P=P, S=S, eta=eta, m=m, symmetrize=True,
                                     threads=cores)
            (lab, it), t_cc = timed(lambda: ogs.cc())
            cpu_extra["cc"] = cpu_block(f"oracle CC on the symmetrised list ({it} iterations)", t_cc, it * 2 * E,
                                        1e9, "GTEPS", cores)
            cpu_extra["cc_symmetrised_build"] = cpu_block("oracle build of the symmetrised list", ogs.build_seconds,
                                                          2 * E, 1e6, "Medges/s", cores)
            del ogs, lab

    if not args.no_engine and args.workload == "rmat24":
        extra["file_load"] = file_load_extras(eg, edges, E, V, P, S, eta, m)

    del g, edges
    if not args.no_engine and not args.no_batching:
        extra["batching"] = batching_extras(eg)
    if not args.no_engine and not args.no_cpu_baseline and args.workload == "rmat24":
        extra["rmat16"] = rmat16_block(eg, cores)
        log(f"[bench] rmat16: pagerank fp64 {extra['rmat16']['pagerank_fp64_gteps']} GTEPS vs cpu "
            f"{extra['rmat16']['cpu_pagerank_fp64']['value']}")
    if not args.no_engine and not args.no_twitter and not args.no_cpu_baseline and args.workload == "rmat24":
        extra["twitter"] = twitter_block(eg, args, cores)

    cpu = None
    if "build" in cpu_extra:
        cb = cpu_extra["build"]
        cpu = {"value": cb["value"], "unit": "Medges/s", "cores": cores, "kind": "port",
               "sample": f"the whole workload: all {E:,} records, same V/P/S/eta/m; oracle/cpu_ref.cpp build "
                         f"(the split_graph-equivalent region) on {cores} threads of a {cpu_extra['host']['model']}, "
                         f"{cb['s']:.2f} s"}
        extra["cpu"] = cpu_extra
        extra["gpu_vs_cpu"] = {"build": round(medges / cb["value"], 1)}
        if "bfs" in cpu_extra:
            extra["gpu_vs_cpu"]["bfs"] = round(extra["bfs"]["gteps"] / cpu_extra["bfs"]["value"], 1)
        if "pagerank_fp64" in cpu_extra:
            extra["gpu_vs_cpu"]["pagerank_fp64"] = round(extra["pagerank"]["gteps_fp64"] /
                                                         cpu_extra["pagerank_fp64"]["value"], 1)
        if "cc" in cpu_extra:
            extra["gpu_vs_cpu"]["cc"] = round(extra["cc"]["gteps"] / cpu_extra["cc"]["value"], 1)

    line = {
        "metric": "CSR+CSC build Medges/s; PageRank GTEPS FP32-vs-FP64 at L1<=1e-6; BFS GTEPS",
        "value": round(medges, 2),
        "unit": "Medges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": desc, "vertices": V, "edges": E, "partitions": P, "sockets": S, "eta": eta,
                   "m": m, "record_bytes": 8, "sort_variant": variant or dict(zip(("scheme", "threads", "digit_bits"),
                                                                                 eg.get_sort_variant(8)))},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
