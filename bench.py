#!/usr/bin/env python3
"""Benchmark of the subgraph-solve hot path on one MI355X.

Default run (what the one JSON line reports):

* value -- the dense kernel pair on an n = 8192 int32 distance matrix (each
  off-diagonal entry an arc with probability 0.01 and weight in [1, 1000],
  else INF, diagonal 0): one blocked Floyd-Warshall closure in place plus
  one min-plus product of two such matrices.  One step = restore the FW
  input (device copy) + FW + min-plus, inputs resident in HBM.  value =
  min-plus GOp/s = (2 n^3 [FW] + 2 n^3 [product]) / step time, 2 ops (add,
  min) per update as the reference counts (fap/kernel/cuda/floyd.cu:340,
  fap/kernel/cuda/minplus.cu:483).
* batch_c4 -- 512 random geometric graphs, n log-uniform in [64, 2048],
  each closed to full APSP: the plain loop (one graph per launch sequence,
  the reference's structure, fap/fap.cpp:189-193) against the batched
  closure replayed from a hipGraph; identical results required.
* serve_c5 -- a Poisson stream at a FIXED arrival rate (--rate req/s): half
  whole-graph APSP requests (RGG, n log-uniform in [256, 4096]), half 1-64
  point queries on the resident 512x512 grid; p50/p99 on the batching
  server and on a plain one (one lane, one request per wave) at the same
  rate; identical results required.
* grid_c3 -- the 512x512 road-like grid, K = 256: one SSSP launch over a
  group's boundary vertices (edges/s, algorithmic GB/s and its fraction of
  HBM), and the plain part-by-part solve against the pipelined groups
  (fapGraph::solve's default) over all 256 parts; identical checksums
  required.

roofline: the Floyd-Warshall closure (fw_diag/fw_panel/fw_rest kernels of
one closure; the largest share of the step), timed with HIP events on its
own stream; roofline_minplus the same for the product kernel.  Min-plus is
not a multiply-accumulate, so the bound is the VALU, not MFMA.  peak = the
measured ceiling of the kernels' instruction mix (one v_add_u32 per update,
one v_min3_i32 per two), probed on the same device in the same run
(fap_probe_minplus_ceiling); on gfx950 v_add_u32 issues at twice the rate of
v_min3_i32, so that ceiling (~76 TOP/s) is above the spec figure of 1 int32
op per lane per clock, 256 CUs x 64 lanes x 2.4 GHz = 39.3 T, which is
reported beside it as spec_peak.

cpu_baseline: the reference's own CPU code (oracle/_ref, compiled from its
headers; floyd_CPU_path and min_plus_path_CPU_advanced, single-threaded like
the reference) on a bounded sample of the step -- the closure of the
leading 1024 x 1024 block and a row band of the product -- each scaled to
the full step by its op count; the int32 restatement (oracle/_build) is used
where the reference build is absent.  host_cores/host_cpu describe the
machine it ran on.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--dense-only]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # spec: 1 int32 op per lane per clock, 39.32 T
# HBM bytes per launch of the n = 8192 min-plus product and per n = 8192
# tile-128 FW closure, from two rocprofv3 counter runs (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, scales from a copy of known size): product 6.47e9 read
# + 0.27e9 written; closure 34.07e9 read + 17.44e9 written.
MINPLUS_8192_TRAFFIC = 6.74e9
FW_8192_T128_TRAFFIC = 51.51e9
# HBM bytes of the SSSP launch over the 512 x 512 grid's parts 1-32 (3,669
# boundary vertices), from the same two counters: 283.4e9 read + 135.4e9
# written.
SSSP_GRID512_G32_TRAFFIC = 418.8e9
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """Logical CPUs, physical cores and model of this host (lscpu)."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        cps, sk = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sk and cps.isdigit() and sk.isdigit():
            info["physical_cores"] = int(cps) * int(sk)
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline_dense(n: int, M0: np.ndarray, A: np.ndarray, B: np.ndarray, rows: int,
                       fw_n: int):
    """GOp/s of the reference CPU path on the dense step, from a sample: the
    FW closure of M0's leading fw_n x fw_n block (scaled by (n / fw_n)^3) and
    `rows` rows of the n^3 product (scaled by n / rows)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    a = np.ascontiguousarray(A[:rows])
    sub = np.ascontiguousarray(M0[:fw_n, :fw_n])
    if O.ref_available():
        t = time.perf_counter()
        O.ref_floyd(sub)
        t_fw = time.perf_counter() - t
        fa_, fb = O.to_ref(a), O.to_ref(B)
        Cm = np.empty((rows, n), np.float32)
        bp = np.zeros((n, n), np.int32)
        cp = np.empty((rows, n), np.int32)
        t = time.perf_counter()
        O.ref().ref_min_plus_cpu_advanced(O._f(fa_), O._f(fb), O._p(bp), O._f(Cm), O._p(cp),
                                          C.c_int32(rows), C.c_int32(n), C.c_int32(n))
        t_mp = time.perf_counter() - t
        kind = "reference"
    else:
        t = time.perf_counter()
        O.floyd(sub)
        t_fw = time.perf_counter() - t
        t = time.perf_counter()
        O.min_plus(a, B)
        t_mp = time.perf_counter() - t
        kind = "port"
    full = t_fw * (n / fw_n) ** 3 + t_mp * (n / rows)
    hi = host_info()
    return {"value": round(4.0 * n ** 3 / full / 1e9, 4), "unit": "GOp/s", "cores": 1,
            "kind": kind,
            "sample": f"floyd_CPU_path on the leading {fw_n}^2 block ({t_fw:.1f} s, x{(n / fw_n) ** 3:.0f}) "
                      f"+ min_plus_path_CPU_advanced on {rows} of {n} rows of the {n}^3 product "
                      f"({t_mp:.1f} s, x{n / rows:.0f}); single-threaded like the reference; "
                      f"scaled step {full:.0f} s",
            "host_cores": hi.get("physical_cores"), "host_logical_cpus": hi.get("logical_cpus"),
            "host_cpu": hi.get("model")}


def dense_section(args, ctx, lib):
    """The c2 step (FW closure + min-plus product at n), timed with HIP
    events per kernel; returns the line's main fields."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    n = args.n
    log(f"[bench] generating n={n} matrices")
    t0 = time.time()
    M0 = synth.dense_matrix(n, 0.01, 1, 1000, 2)
    A = synth.dense_matrix(n, 0.01, 1, 1000, 21)
    B = synth.dense_matrix(n, 0.01, 1, 1000, 22)
    log(f"[bench] generated in {time.time() - t0:.1f}s")
    ceiling = ctx.probe_minplus_ceiling()
    log(f"[bench] min-plus instruction-mix ceiling {ceiling:.2f} TOP/s")
    tuned = _tune(ctx, args, "fw", fa.TuneCache.key_fw(n), lambda: ctx.tune_fw(n, reps=1))
    d_src = ctx.to_device(M0)
    d_fw = ctx.alloc((n, n))
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    dC = ctx.alloc((n, n))
    s = ctx.stream
    nev = 4 * (args.steps + args.warmup)
    ev = [C.c_void_p() for _ in range(nev)]
    for e in ev:
        fa.check(lib.fap_event_create(C.byref(e)))

    def step(i):
        fa.check(lib.fap_copy_i32(ctx.handle, d_fw.ptr, d_src.ptr, n * n, s))
        fa.check(lib.fap_event_record(ev[4 * i], s))
        ctx.fw(d_fw.ptr, n, n, s)
        fa.check(lib.fap_event_record(ev[4 * i + 1], s))
        fa.check(lib.fap_event_record(ev[4 * i + 2], s))
        ctx.minplus(dA.ptr, n, dB.ptr, n, dC.ptr, n, n, n, n, False, s)
        fa.check(lib.fap_event_record(ev[4 * i + 3], s))

    for i in range(args.warmup):
        step(args.steps + i)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    ctx.sync()
    total = time.perf_counter() - t0
    fw_ms, mp_ms = [], []
    for i in range(args.steps):
        ms = C.c_float()
        fa.check(lib.fap_event_elapsed(ev[4 * i], ev[4 * i + 1], C.byref(ms)))
        fw_ms.append(ms.value)
        fa.check(lib.fap_event_elapsed(ev[4 * i + 2], ev[4 * i + 3], C.byref(ms)))
        mp_ms.append(ms.value)
        log(f"[bench] step {i}: fw {fw_ms[-1]:.2f} ms, min-plus {mp_ms[-1]:.2f} ms")
    for e in ev:
        lib.fap_event_destroy(e)
    for d in (d_src, d_fw, dA, dB, dC):
        d.free()
    ops_step = 4.0 * n ** 3
    mp_avg = float(np.mean(mp_ms))
    fw_avg = float(np.mean(fw_ms))
    ach = 2.0 * n ** 3 / (mp_avg * 1e-3) / 1e12
    ach_fw = 2.0 * n ** 3 / (fw_avg * 1e-3) / 1e12
    res = {
        "metric": "min-plus GOp/s",
        "value": round(ops_step / (total / args.steps) / 1e9, 1),
        "unit": "GOp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"dense FW closure + min-plus product, n={n}, arc p=0.01, "
                               f"w 1..1000 (value); batch_c4, serve_c5 and grid_c3 carry the "
                               f"batched/served-vs-plain forms", "n": n, "ops_per_step": ops_step},
        "roofline": {"bound": "valu",
                     "kernel": "Floyd-Warshall closure (fw_rest phase-3 tiles with the fused "
                               "next diagonal, fw_panel, fw_diag; one closure per step)",
                     "achieved": round(ach_fw, 2), "peak": round(ceiling, 2), "unit": "TOP/s",
                     "frac": round(ach_fw / ceiling, 3),
                     "traffic": (FW_8192_T128_TRAFFIC if n == 8192 and tuned
                                 and tuned.get("params", [0])[0] == 128 else None),
                     "peak_kind": "measured v_add_u32 + v_min3_i32 mix ceiling",
                     "spec_peak": round(VALU_PEAK_TOPS, 2),
                     "frac_of_spec_peak": round(ach_fw / VALU_PEAK_TOPS, 3),
                     "ms_per_launch": round(fw_avg, 3), "ops_per_launch": 2.0 * n ** 3},
        "roofline_minplus": {"bound": "valu", "kernel": "minplus_gemm_kernel",
                             "achieved": round(ach, 2), "peak": round(ceiling, 2),
                             "unit": "TOP/s", "frac": round(ach / ceiling, 3),
                             "traffic": MINPLUS_8192_TRAFFIC if n == 8192 else None,
                             "spec_peak": round(VALU_PEAK_TOPS, 2),
                             "frac_of_spec_peak": round(ach / VALU_PEAK_TOPS, 3),
                             "ms_per_launch": round(mp_avg, 3), "ops_per_launch": 2.0 * n ** 3},
        "tuned": tuned,
    }
    res["cpu_baseline"] = None
    if args.cpu_rows > 0:
        log("[bench] cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline_dense(n, M0, A, B, rows=args.cpu_rows,
                                                 fw_n=args.cpu_fw_n)
    return res, ceiling


def bench_dense(args):
    """The driver's line: the dense step as value, plus the batched and
    served forms against their plain counterparts (c4, c5, c3)."""
    import fastapsp_amd as fa
    lib = fa.load()
    t_all = time.time()
    ctx = fa.Context(0)
    res, ceiling = dense_section(args, ctx, lib)
    rc = 0
    if not args.dense_only:
        t0 = time.time()
        res["batch_c4"] = batch_section(args, ctx, lib, ceiling)
        log(f"[bench] batch_c4 in {time.time() - t0:.1f} s: {res['batch_c4']}")
        rc |= 0 if res["batch_c4"]["identical"] else 1
        ctx.close()
        grid = grid_setup(args)
        t0 = time.time()
        res["serve_c5"] = serve_section(args, grid)
        log(f"[bench] serve_c5 in {time.time() - t0:.1f} s: {res['serve_c5']}")
        rc |= 0 if res["serve_c5"]["identical"] else 1
        t0 = time.time()
        res["grid_c3"] = grid_section(args, grid, modes=("plain", "single", "pipelined"))
        log(f"[bench] grid_c3 in {time.time() - t0:.1f} s: {res['grid_c3']}")
        rc |= 0 if res["grid_c3"]["identical"] else 1
    else:
        ctx.close()
    res["bench_wall_s"] = round(time.time() - t_all, 1)
    print(json.dumps(res), flush=True)
    return rc


def _tune(ctx, args, kind, key, run):
    """Attach a tune cache (--tune-cache file if given), measure the bucket if
    it has no entry, save the file back; returns the chosen parameters."""
    import os
    import fastapsp_amd as fa
    path = args.tune_cache
    cache = fa.TuneCache(path if path and os.path.exists(path) else None)
    ctx.set_tune_cache(cache)
    hit = cache.get(kind, key)
    if hit is None:
        t0 = time.perf_counter()
        run()
        hit = cache.get(kind, key)
        log(f"[bench] tuned {kind} bucket {key} in {time.perf_counter() - t0:.2f}s: {hit}")
        if path:
            cache.save(path)
    return {"kind": kind, "bucket": int(key), "params": list(hit[0]), "ms": round(hit[1], 3)}


def bench_copy(args):
    """Streaming copy of n*n int32 (known bytes): HBM reference and the
    calibration kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import fastapsp_amd as fa
    lib = fa.load()
    n = args.n
    ctx = fa.Context(0)
    a, b = ctx.alloc((n, n)), ctx.alloc((n, n))
    ctx.fill(a, 1)
    s = ctx.stream
    ev = [C.c_void_p() for _ in range(2)]
    for e in ev:
        fa.check(lib.fap_event_create(C.byref(e)))
    for _ in range(args.warmup):
        fa.check(lib.fap_copy_i32(ctx.handle, b.ptr, a.ptr, n * n, s))
    ctx.sync()
    fa.check(lib.fap_event_record(ev[0], s))
    for _ in range(args.steps):
        fa.check(lib.fap_copy_i32(ctx.handle, b.ptr, a.ptr, n * n, s))
    fa.check(lib.fap_event_record(ev[1], s))
    ms = C.c_float()
    fa.check(lib.fap_event_elapsed(ev[0], ev[1], C.byref(ms)))
    per = ms.value / args.steps
    gbs = 2 * 4 * n * n / (per * 1e-3) / 1e9
    print(json.dumps({"metric": "copy GB/s", "value": round(gbs, 1), "unit": "GB/s",
                      "bytes_per_launch": 8 * n * n, "ms_per_launch": round(per, 4),
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "higher_is_better": True, "dtype": "int32", "data": "synthetic",
                      "config": {"workload": f"copy {n}x{n} int32"}}), flush=True)
    return 0


def bench_grid_pipeline(args):
    """Full pipeline, plain form: 512x512 road-like grid (10% of edges removed,
    2% of V diagonal shortcuts, w 1..100), K=256, parts solved one after the
    other through the device stages (SSSP -> gather -> FW -> min-plus ->
    scatter), each 1024 x 262144 row block checksummed on the device."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    lib = fa.load()
    R, K = args.grid, args.parts_k
    V = R * R
    u, v, w = synth.grid_graph(R, R, 0.10, 0.02, 1, 100, 3)
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    E = len(col)
    g = fa.Graph.from_csr(ro, col, wt.astype(np.float32), K)
    g.preCondition()
    vn, vo, bn, _ = g.part_info()
    st2ed = g.getMapping()
    gid = g.getGraphId()
    ed2st = np.empty(V, np.int32)
    ed2st[st2ed] = np.arange(V, dtype=np.int32)
    g.close()
    ctx = fa.Context(0)
    dg = ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, st2ed, ed2st)
    nmax = int(vn[1:].max())
    blk = ctx.alloc((nmax, V))
    ld_sub = (nmax + 3) // 4 * 4
    sub = ctx.alloc((nmax, ld_sub))
    d_st2ed = ctx.to_device(st2ed)
    s = ctx.stream
    parts = list(range(1, K + 1))[: args.max_parts] if args.max_parts else list(range(1, K + 1))
    nev = 5
    ev = [C.c_void_p() for _ in range(nev)]
    for e in ev:
        fa.check(lib.fap_event_create(C.byref(e)))
    t_sssp = t_fw = t_mp = t_glue = 0.0
    ops_fw = ops_mp = 0.0
    sums = []
    ms = C.c_float()

    def part(pid):
        st, n, b = int(vo[pid]), int(vn[pid]), int(bn[pid])
        ldsb = (n + 3) // 4 * 4
        fa.check(lib.fap_event_record(ev[0], s))
        dg.sssp(d_st2ed.offset(st), b, blk.ptr, V, s)
        fa.check(lib.fap_event_record(ev[1], s))
        dg.gather(st, n, b, blk.ptr, V, sub.ptr, ldsb, s)
        ctx.fw(sub.ptr, n, ldsb, s)
        fa.check(lib.fap_event_record(ev[2], s))
        ctx.minplus(sub.offset(b * ldsb), ldsb, blk.ptr, V, blk.offset(b * V), V, n - b, V, b,
                    False, s)
        fa.check(lib.fap_event_record(ev[3], s))
        dg.scatter(st, n, n, sub.ptr, ldsb, blk.ptr, V, s)
        fa.check(lib.fap_event_record(ev[4], s))
        return n, b

    part(parts[0])  # warm-up
    ctx.sync()
    t0 = time.perf_counter()
    for pid in parts:
        n, b = part(pid)
        sums.append(ctx.checksum(blk.ptr, n * V))  # synchronizes
        fa.check(lib.fap_event_elapsed(ev[0], ev[1], C.byref(ms)))
        t_sssp += ms.value
        fa.check(lib.fap_event_elapsed(ev[1], ev[2], C.byref(ms)))
        t_fw += ms.value
        fa.check(lib.fap_event_elapsed(ev[2], ev[3], C.byref(ms)))
        t_mp += ms.value
        fa.check(lib.fap_event_elapsed(ev[3], ev[4], C.byref(ms)))
        t_glue += ms.value
        ops_fw += 2.0 * n ** 3
        ops_mp += 2.0 * (n - b) * b * V
    total = time.perf_counter() - t0
    P = len(parts)
    srcs = float(sum(int(bn[p]) for p in parts))
    sssp_bytes = P * (4.0 * (V + 1) + 8.0 * E) + 4.0 * srcs * V
    res = {
        "metric": "min-plus GOp/s", "value": round((ops_fw + ops_mp) / total / 1e9, 1),
        "unit": "GOp/s", "n_gpus": 1, "steps": P, "warmup": 1,
        "ms_per_step": round(1e3 * total / P, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{R}x{R} grid (-10% edges, +2% diagonals, w 1..100), K={K}, "
                               f"plain part-by-part solve", "V": V, "E": E, "K": K,
                   "parts_timed": P, "mean_part": float(vn[1:].mean()),
                   "mean_boundary": float(bn[1:].mean())},
        "stage_ms": {"sssp": round(t_sssp, 2), "gather+fw": round(t_fw, 2),
                     "minplus": round(t_mp, 2), "scatter": round(t_glue, 2),
                     "wall_incl_checksum": round(total * 1e3, 2)},
        "minplus_stage": {"ops": ops_mp, "TOP/s": round(ops_mp / (t_mp * 1e-3) / 1e12, 2),
                          "frac_of_spec_peak": round(ops_mp / (t_mp * 1e-3) / 1e12 / VALU_PEAK_TOPS, 3)},
        "fw_stage": {"ops": ops_fw, "TOP/s": round(ops_fw / (t_fw * 1e-3) / 1e12, 2)},
        "sssp_stage": {"sources": srcs, "edges_per_s": srcs * E / (t_sssp * 1e-3),
                       "algorithmic_GB_per_s": round(sssp_bytes / (t_sssp * 1e-3) / 1e9, 1)},
        "checksum_first_parts": [list(map(int, x)) for x in sums[:3]],
    }
    print(json.dumps(res), flush=True)
    return 0


def grid_setup(args):
    """The 512x512 road-like grid (10% of edges removed, 2% of V diagonal
    shortcuts, w 1..100) and its K-part partition (host preCondition)."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    R, K = args.grid, args.parts_k
    V = R * R
    t0 = time.time()
    u, v, w = synth.grid_graph(R, R, 0.10, 0.02, 1, 100, 3)
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    g = fa.Graph.from_csr(ro, col, wt.astype(np.float32), K)
    g.preCondition()
    vn, vo, bn, _ = g.part_info()
    st2ed = g.getMapping()
    gid = g.getGraphId()
    g.close()
    ed2st = np.empty(V, np.int32)
    ed2st[st2ed] = np.arange(V, dtype=np.int32)
    log(f"[bench] grid {R}x{R} K={K}: V={V} E={len(col)}, partitioned in {time.time() - t0:.1f} s")
    return dict(R=R, K=K, V=V, E=len(col), ro=ro, col=col, wt=wt, vn=vn, vo=vo, bn=bn,
                st2ed=st2ed, gid=gid, ed2st=ed2st)


def grid_section(args, grid, modes=("plain", "pipelined")):
    """c3 on the grid: (1) one SSSP launch over the boundary vertices of the
    first --group parts (edges/s, algorithmic bytes 4(V+1) + 8E + 4 S V per
    launch, their rate as a fraction of HBM peak), and the same rows through
    the boundary index (fap_subgraph_boundary_rows); (2) all parts solved
    plain (fap_subgraph_solve part after part: SSSP stage 1, the reference's
    pipeline), single (one fap_subgraph_solve_batch per part, stage 1 through
    the boundary index -- fapGraph::solveSubGraph's default), batched (groups
    of --group parts, fap_subgraph_solve_batch) and/or pipelined (groups alternating
    over two streams, fap_subgraph_solve_groups -- fapGraph::solve's
    default), in calls of --call-parts parts whose ~1.1 GiB row blocks stay
    allocated; row blocks are checksummed on the device after each call
    (outside the timed region) and must agree between the forms."""
    import fastapsp_amd as fa
    lib = fa.load()
    V, E, K, G = grid["V"], grid["E"], grid["K"], args.group
    vn, vo, bn, st2ed = grid["vn"], grid["vo"], grid["bn"], grid["st2ed"]
    ctx = fa.Context(0)
    dg = ctx.upload_graph(grid["ro"], grid["col"], grid["wt"])
    dg.set_partition(grid["gid"], st2ed, grid["ed2st"])
    parts = list(range(1, K + 1))[: args.max_parts] if args.max_parts else list(range(1, K + 1))
    srcs = np.concatenate([st2ed[int(vo[p]): int(vo[p]) + int(bn[p])]
                           for p in parts[:G]]).astype(np.int32)
    tuned = _tune(ctx, args, "sssp", fa.TuneCache.key_sssp(V, E, len(srcs)),
                  lambda: ctx.tune_sssp(dg, srcs, reps=1))
    dsrc = ctx.to_device(srcs)
    out = ctx.alloc((len(srcs), V))
    sssp_ms = _time_ms(lib, ctx, lambda: dg.sssp(dsrc.ptr, len(srcs), out.ptr, V, ctx.stream),
                       reps=3)
    out.free()
    dsrc.free()
    S = len(srcs)
    sssp_bytes = 4.0 * (V + 1) + 8.0 * E + 4.0 * S * V
    sssp = {"sources": S, "parts": min(G, len(parts)), "ms_per_launch": round(sssp_ms, 3),
            "edges_per_s": round(S * E / (sssp_ms * 1e-3), 1),
            "algorithmic_bytes": sssp_bytes,
            "algorithmic_GB_per_s": round(sssp_bytes / (sssp_ms * 1e-3) / 1e9, 2),
            "frac_of_hbm_peak": round(sssp_bytes / (sssp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "tuned": tuned}
    if args.grid == 512 and S == 3669:   # the counted launch: its measured HBM traffic
        sssp["traffic"] = SSSP_GRID512_G32_TRAFFIC
        sssp["traffic_GB_per_s"] = round(SSSP_GRID512_G32_TRAFFIC / (sssp_ms * 1e-3) / 1e9, 1)
        sssp["traffic_frac_of_hbm_peak"] = round(sssp["traffic_GB_per_s"] / HBM_PEAK_GBS, 3)
    # the same sources' rows through the boundary index (stage 1 of the
    # batched forms), into the first parts' row blocks
    grp = parts[:G]
    rbs = [ctx.alloc((int(bn[p]), V)) for p in grp]
    bargs = ([int(vo[p]) for p in grp], [int(vn[p]) for p in grp], [int(bn[p]) for p in grp],
             [b.ptr for b in rbs])
    dg.boundary_rows(*bargs, V)          # builds the index (timed below without it)
    idx_ms = _time_ms(lib, ctx, lambda: dg.boundary_rows(*bargs, V), reps=3)
    for b in rbs:
        b.free()
    sssp["index_ms_per_call"] = round(idx_ms, 3)
    sssp["index_edges_per_s_equiv"] = round(S * E / (idx_ms * 1e-3), 1)  # same rows, S.E / t
    sssp["index_over_sssp"] = round(sssp_ms / idx_ms, 2)
    log(f"[bench] c3 SSSP: {sssp}")
    P = max(G, min(args.call_parts, len(parts)))
    calls = [parts[i:i + P] for i in range(0, len(parts), P)]
    nmax = int(vn[1:].max())
    blks = [ctx.alloc((nmax, V)) for _ in range(min(P, len(parts)))]

    def args_of(ids):
        return ([int(vo[p]) for p in ids], [int(vn[p]) for p in ids], [int(bn[p]) for p in ids],
                [blks[i].ptr for i in range(len(ids))])

    def run(mode):
        t = 0.0
        sums = []
        for call in calls:
            ctx.sync()
            t0 = time.perf_counter()
            if mode == "plain":
                for i, p in enumerate(call):
                    dg.subgraph_solve(int(vo[p]), int(vn[p]), int(bn[p]), blks[i].ptr, V)
            elif mode == "single":
                for i, p in enumerate(call):
                    dg.subgraph_solve_batch([int(vo[p])], [int(vn[p])], [int(bn[p])],
                                            [blks[i].ptr], V)
            elif mode == "batched":
                for j in range(0, len(call), G):
                    st, ns, bs, _ = args_of(call[j:j + G])
                    dg.subgraph_solve_batch(st, ns, bs, [blks[j + i].ptr for i in range(len(st))],
                                            V)
            else:
                st, ns, bs, rb = args_of(call)
                dg.subgraph_solve_groups(st, ns, bs, rb, V, G)
            ctx.sync()
            t += time.perf_counter() - t0
            sums += [ctx.checksum(blks[i].ptr, int(vn[p]) * V) for i, p in enumerate(call)]
        return t, sums

    for m in modes:
        run(m)  # warm-up (workspaces, kernels, the second stream's context)
    res = {}
    for mode in modes:
        ts = []
        for _ in range(max(1, args.grid_steps)):
            t, sums = run(mode)
            ts.append(t)
        res[mode] = (min(ts), sums)
        log(f"[bench] c3 {mode}: {1e3 * min(ts):.1f} ms for {len(parts)} parts")
    for b in blks:
        b.free()
    dg.close()
    ctx.close()
    ops = sum(2.0 * int(vn[p]) ** 3 + 2.0 * (int(vn[p]) - int(bn[p])) * int(bn[p]) * V
              for p in parts)
    sums0 = res[modes[0]][1]
    out = {"workload": f"{grid['R']}x{grid['R']} grid (-10% edges, +2% diagonals, w 1..100), "
                       f"K={K}, every part's n_sub x V row block",
           "V": V, "E": E, "K": K, "parts": len(parts), "group": G, "call_parts": P,
           "ops": ops, "sssp": sssp}
    for mode in modes:
        out[mode + "_ms"] = round(1e3 * res[mode][0], 2)
        out[mode + "_GOp_s"] = round(ops / res[mode][0] / 1e9, 1)
    for mode in modes[1:]:
        out[mode + "_over_plain"] = round(res["plain"][0] / res[mode][0], 2)
    out["identical"] = all(res[m][1] == sums0 for m in modes)
    return out


def bench_grid_batch(args):
    """c3 alone: plain vs batched vs pipelined per-part solve on the grid."""
    out = grid_section(args, grid_setup(args), modes=("plain", "single", "batched", "pipelined"))
    res = {"metric": "min-plus GOp/s", "value": out["pipelined_GOp_s"], "unit": "GOp/s",
           "n_gpus": 1, "steps": args.grid_steps, "warmup": 1,
           "ms_per_step": out["pipelined_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32", "data": "synthetic",
           "config": {"workload": out["workload"] + ", pipelined groups"}, "grid_c3": out}
    print(json.dumps(res), flush=True)
    return 0 if out["identical"] else 1


def _serve_pool(count, V_res, seed):
    """Request mix for the served workload: half whole-graph solves (random
    geometric graphs, n log-uniform in [256, 4096], average degree 6, directed,
    w 1..255), half batches of 1-64 (src, dst) queries on the resident graph."""
    from fastapsp_amd import synth
    rng = np.random.default_rng(seed)
    kinds = rng.random(count) < 0.5
    sizes = synth.loguniform_sizes(count, 256, 4096, seed)
    pool = []
    for i in range(count):
        if kinds[i]:
            n = int(sizes[i])
            u, v, w = synth.rgg_graph(n, 6, 1, 255, 1000 + i)
            pool.append(("graph", synth.to_csr(n, u, v, w, True)))
        else:
            k = int(rng.integers(1, 65))
            pool.append(("queries", (rng.integers(0, V_res, k).astype(np.int32),
                                     rng.integers(0, V_res, k).astype(np.int32))))
    return pool


def _serve_run(srv, pool, arrivals):
    """Submits pool[i] at arrivals[i] seconds (None: all at once) from one
    thread while this thread collects results in order; returns per-request
    records (server-clock latencies) and the wall time."""
    import threading
    ids = [None] * len(pool)
    ready = threading.Semaphore(0)

    def submitter():
        t0 = time.perf_counter()
        for i, (kind, payload) in enumerate(pool):
            if arrivals is not None:
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            ids[i] = srv.submit_graph(*payload) if kind == "graph" else srv.submit_queries(*payload)
            ready.release()

    th = threading.Thread(target=submitter)
    t_start = time.perf_counter()
    th.start()
    recs = []
    for i, (kind, payload) in enumerate(pool):
        ready.acquire()
        info = srv.wait(ids[i])
        out = srv.checksum(ids[i]) if kind == "graph" else tuple(srv.answers(ids[i]).tolist())
        srv.release(ids[i])
        recs.append((kind, info, out))
    th.join()
    wall = time.perf_counter() - t_start
    return recs, wall


def _serve_summary(recs, wall, pool):
    lat = np.array([r[1]["latency_us"] for r in recs]) / 1e3
    lg = np.array([r[1]["latency_us"] for r in recs if r[0] == "graph"]) / 1e3
    lq = np.array([r[1]["latency_us"] for r in recs if r[0] == "queries"]) / 1e3
    ops = sum(2.0 * (len(p[1][0]) - 1) ** 3 for p in pool if p[0] == "graph")
    first = min(r[1]["submit_us"] for r in recs)
    last = max(r[1]["done_us"] for r in recs)
    span = (last - first) / 1e6

    def pct(a, q):
        return round(float(np.percentile(a, q)), 3) if a.size else None

    return {"requests": len(recs), "span_s": round(span, 3), "req_per_s": round(len(recs) / span, 1),
            "GOp_s": round(ops / span / 1e9, 1),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "graph_p50_ms": pct(lg, 50), "graph_p99_ms": pct(lg, 99),
            "query_p50_ms": pct(lq, 50), "query_p99_ms": pct(lq, 99),
            "mean_wave": round(float(np.mean([r[1]["wave_size"] for r in recs])), 2)}


def serve_section(args, grid):
    """c5: served solves at a fixed arrival rate.  A burst of the whole mix
    first measures the saturation rate (reported, not used); then the mix
    arrives as a Poisson stream at --rate req/s on the continuous-batching
    server (queries through the index over --query-parts parts) and on a
    plain one (one lane, one request per wave, an SSSP per query batch).
    Results of the three runs must be identical."""
    import fastapsp_amd as fa
    ro, col, wt, V = grid["ro"], grid["col"], grid["wt"], grid["V"]
    pool = _serve_pool(args.requests, V, 5)
    log(f"[bench] serve: {len(pool)} requests ({sum(p[0] == 'graph' for p in pool)} graphs), "
        f"resident V={V}, rate {args.rate} req/s")
    budget = int(args.hbm_gib * (1 << 30))
    lam = float(args.rate)
    gaps = np.random.default_rng(11).exponential(1.0 / lam, len(pool))
    arrivals = np.cumsum(gaps) - gaps[0]
    with fa.Server(hbm_budget=budget, lanes=args.lanes, query_parts=args.query_parts) as srv:
        t0 = time.perf_counter()
        srv.attach_graph(ro, col, wt)
        attach_s = time.perf_counter() - t0
        ist = srv.stats()
        log(f"[bench]   attach {attach_s:.2f} s (index {ist['index_bytes'] / 2**30:.1f} GiB, "
            f"{ist['index_build_ms']:.0f} ms)")
        _serve_run(srv, pool[:16], None)  # warm-up
        recs, wall = _serve_run(srv, pool, None)
        sat = _serve_summary(recs, wall, pool)
        ref = [r[2] for r in recs]
        recs, wall = _serve_run(srv, pool, arrivals)
        served = _serve_summary(recs, wall, pool)
        same = [r[2] for r in recs] == ref
        st = srv.stats()
    log(f"[bench]   served: {served}")
    with fa.Server(hbm_budget=budget, lanes=1, max_batch=1) as srv:
        srv.attach_graph(ro, col, wt)
        _serve_run(srv, pool[:16], None)
        recs, wall = _serve_run(srv, pool, arrivals)
        plain = _serve_summary(recs, wall, pool)
        same = same and [r[2] for r in recs] == ref
    log(f"[bench]   plain: {plain}")
    return {"workload": f"Poisson stream at {lam:g} req/s, half RGG APSP (n logU[256,4096], "
                        f"deg 6, w 1..255), half 1-64 point queries on the {grid['R']}x"
                        f"{grid['R']} grid (V={V}); HBM budget {args.hbm_gib} GiB, "
                        f"{args.lanes} lanes",
            "arrival_rate_req_s": lam, "requests": len(pool),
            "p50_ms": served["p50_ms"], "p99_ms": served["p99_ms"],
            "plain_p50_ms": plain["p50_ms"], "plain_p99_ms": plain["p99_ms"],
            "served_over_plain_p99": round(plain["p99_ms"] / served["p99_ms"], 2),
            "served": served, "plain_same_rate": plain, "saturation_burst": sat,
            "identical": bool(same),
            "server_stats": {k: st[k] for k in ("waves", "max_wave", "arena_peak", "index_bytes",
                                                "index_build_ms")},
            "query_index": {"parts": args.query_parts, "attach_s": round(attach_s, 3)}}


def bench_serve(args):
    """c5 alone."""
    out = serve_section(args, grid_setup(args))
    res = {"metric": "served p50/p99 solve latency at a fixed arrival rate", "value": out["p99_ms"],
           "unit": "ms (p99)", "n_gpus": 1, "steps": out["requests"], "warmup": 16,
           "ms_per_step": round(1e3 * out["served"]["span_s"] / out["requests"], 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic", "config": {"workload": out["workload"]}, "serve_c5": out}
    print(json.dumps(res), flush=True)
    return 0 if out["identical"] else 1


def bench_c1(args):
    """Plain method, small: 64x64 grid (5% of edges removed), unit weights,
    K=8, the full 4096 x 4096 matrix through the drop-in fapGraph::solve(),
    plain (parts one after another) and batched; the reference's own CPU
    pipeline (oracle/_ref, Dijkstra over OpenMP threads, serial FW and
    min-plus) timed on all 8 parts as the baseline."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    R, K = 64, 8
    u, v, w = synth.grid_graph(R, R, 0.05, 0.0, 1, 1, 1)
    V = R * R
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    res = {}
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        g.preCondition()
        vn, vo, bn, _ = g.part_info()
        ops = sum(2.0 * int(vn[p]) ** 3 + 2.0 * (int(vn[p]) - int(bn[p])) * int(bn[p]) * V
                  for p in range(1, K + 1))
        for mode in ("plain", "batched"):
            g.setBatched(mode == "batched")
            g.solve()
            ts = []
            for _ in range(max(args.steps, 1)):
                t0 = time.perf_counter()
                g.solve()
                ts.append(time.perf_counter() - t0)
            res[mode] = min(ts)
            res[mode + "_sum"] = sum(g.checksum(p)[1] for p in range(1, K + 1)) % (1 << 64)
            log(f"[c1] {mode}: {1e3 * res[mode]:.2f} ms")
        gid = g.getGraphId()
        pc = {"st2ed": g.getMapping().astype(np.int32), "ver_num": vn.astype(np.int32),
              "ver_off": vo.astype(np.int32), "bdy_num": bn.astype(np.int32)}
        pc["ed2st"] = np.empty(V, np.int32)
        pc["ed2st"][pc["st2ed"]] = np.arange(V, dtype=np.int32)
        D = g.getAllDistanceInt32()
    cpu = None
    if args.cpu_rows > 0:
        t0 = time.perf_counter()
        rows = []
        kind = "reference" if O.ref_available() else "port"
        for p in range(1, K + 1):
            if kind == "reference":
                rows.append(O.from_ref(O.ref_run_part(p, ro, col, wt, gid, pc)))
            else:
                rows.append(O.run_part(p, ro, col, wt, gid, pc))
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(np.concatenate(rows), D))
        cpu = {"value": round(ops / dt / 1e9, 3), "unit": "GOp/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())) if kind == "reference" else 1,
               "kind": kind,
               "sample": f"all {K} parts through the reference's CPU run() ({dt:.2f} s); "
                         f"rows identical to the GPU's: {same}"}
    out = {
        "metric": "min-plus GOp/s", "value": round(ops / res["batched"] / 1e9, 1), "unit": "GOp/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(1e3 * res["batched"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "64x64 grid (-5% edges), unit weights, K=8, full 4096^2 matrix "
                               "via fapGraph::solve()", "V": V, "K": K, "ops": ops},
        "plain_ms": round(1e3 * res["plain"], 3), "batched_ms": round(1e3 * res["batched"], 3),
        "batched_over_plain": round(res["plain"] / res["batched"], 2),
        "identical": res["plain_sum"] == res["batched_sum"], "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def _time_ms(lib, ctx, fn, reps=2):
    """Average device milliseconds of fn() on the context stream."""
    import fastapsp_amd as fa
    e0, e1 = C.c_void_p(), C.c_void_p()
    fa.check(lib.fap_event_create(C.byref(e0)))
    fa.check(lib.fap_event_create(C.byref(e1)))
    fn()
    ctx.sync()
    fa.check(lib.fap_event_record(e0, ctx.stream))
    for _ in range(reps):
        fn()
    fa.check(lib.fap_event_record(e1, ctx.stream))
    ms = C.c_float()
    fa.check(lib.fap_event_elapsed(e0, e1, C.byref(ms)))
    lib.fap_event_destroy(e0)
    lib.fap_event_destroy(e1)
    return ms.value / reps


def bench_fw_sweep(args):
    """FW closure time per (n, block size, look-ahead): the tuning table."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    lib = fa.load()
    ctx = fa.Context(0)
    rows = []
    for n in [int(x) for x in args.sizes.split(",")]:
        m = synth.dense_matrix(n, min(1.0, 6.0 / n) if n < 600 else 0.01, 1, 1000, n)
        src = ctx.to_device(m)
        d = ctx.alloc((n, n))
        for tile in (64, 128):
            for look in (0, 1):
                ctx.set_option("fw_tile", tile)
                ctx.set_option("fw_lookahead", look)

                def run():
                    fa.check(lib.fap_copy_i32(ctx.handle, d.ptr, src.ptr, n * n, ctx.stream))
                    ctx.fw(d.ptr, n, n, ctx.stream)
                copy_ms = _time_ms(lib, ctx, lambda: fa.check(
                    lib.fap_copy_i32(ctx.handle, d.ptr, src.ptr, n * n, ctx.stream)))
                ms = _time_ms(lib, ctx, run, reps=args.steps) - copy_ms
                r = {"n": n, "tile": tile, "lookahead": look, "ms": round(ms, 4),
                     "TOP/s": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 3)}
                rows.append(r)
                print(json.dumps(r), flush=True)
        src.free()
        d.free()
    ctx.set_option("fw_tile", 0)
    ctx.set_option("fw_lookahead", 0)
    return 0


def bench_sssp_sweep(args):
    """Multi-source SSSP time per (threads, workgroups per CU, state layout,
    queue binning) on the 512x512 grid (w 1..100, K=256 partition): the boundary vertices of one
    part (plain per-part stage 1) and of 32 parts (a batched group)."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    lib = fa.load()
    ctx = fa.Context(0)
    R, K = args.grid, args.parts_k
    V = R * R
    u, v, w = synth.grid_graph(R, R, 0.10, 0.02, 1, 100, 3)
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    g = fa.Graph.from_csr(ro, col, wt.astype(np.float32), K)
    g.preCondition()
    vn, vo, bn, _ = g.part_info()
    st2ed = g.getMapping()
    g.close()
    dg = ctx.upload_graph(ro, col, wt)
    for parts in (1, 32):
        srcs = np.concatenate([st2ed[int(vo[p]): int(vo[p]) + int(bn[p])]
                               for p in range(1, parts + 1)]).astype(np.int32)
        dsrc = ctx.to_device(srcs)
        out = ctx.alloc((len(srcs), V))
        ref = None
        for threads, wgs in ((256, 4), (256, 8), (512, 2), (1024, 1)):
            for pack, binq, arcs in ((1, 3, 1), (1, 3, 2), (1, 3, 4), (1, 0, 2), (2, 3, 2)):
                ctx.set_option("sssp_threads", threads)
                ctx.set_option("sssp_wgs", wgs)
                ctx.set_option("sssp_pack", pack)
                ctx.set_option("sssp_bin", binq)
                ctx.set_option("sssp_arcs", arcs)
                ms = _time_ms(lib, ctx, lambda: dg.sssp(dsrc.ptr, len(srcs), out.ptr, V,
                                                        ctx.stream), reps=1)
                cs = ctx.checksum(out.ptr, len(srcs) * V)
                ref = cs if ref is None else ref
                byts = 4.0 * (V + 1) + 8.0 * len(col) + 4.0 * len(srcs) * V
                r = {"graph": f"{R}x{R} w1..100", "parts": parts, "sources": len(srcs),
                     "threads": threads, "wgs_per_cu": wgs, "pack": pack, "bin": binq,
                     "arcs": arcs,
                     "ms": round(ms, 3),
                     "edges_per_s": len(srcs) * len(col) / (ms * 1e-3),
                     "algorithmic_GB_per_s": round(byts / (ms * 1e-3) / 1e9, 1),
                     "same_rows": cs == ref}
                print(json.dumps(r), flush=True)
        out.free()
    dg.close()
    ctx.set_option("sssp_threads", 0)
    ctx.set_option("sssp_wgs", 0)
    ctx.set_option("sssp_pack", 0)
    ctx.set_option("sssp_arcs", 0)
    ctx.set_option("sssp_bin", 3)
    return 0


def batch_section(args, ctx, lib, ceiling):
    """c4: 512 random geometric graphs, n log-uniform in [64, 2048] (points in
    the unit square, radius for average degree 6, both arcs with independent
    weights 1..255), each closed to full APSP as one part (K = 1: the
    reference's pipeline with no boundary vertices, so its SSSP stage is
    empty -- dense build + blocked FW).  Plain = one graph per launch
    sequence, looped on the host (fap/fap.cpp:189-193); batched = one
    phase-synchronous launch sequence over all graphs (size-ordered task
    queues, paired pivot rounds, two chains), replayed from a hipGraph;
    persistent = every phase of every graph in ONE launch whose workgroups
    claim tasks in queue order and wait in the kernel for their graph's
    earlier phases (a persistent plan, fap_fw_plan_create_persistent:
    queues built once, each run a replayed hipGraph of a counter reset and
    the launch; fap_fw_plan_check after the timed runs).  Rows are 128-byte aligned (ld a multiple of 32), as
    the single launch requires.  Checksums of all matrices must agree."""
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    sizes = synth.loguniform_sizes(args.jobs, 64, 2048, 4).astype(int)
    graphs, ptrs, lds = [], [], []
    ld_of = [(n + 31) // 32 * 32 for n in sizes]
    total = sum(n * ld for n, ld in zip(sizes, ld_of))
    pool = ctx.alloc(total)
    off = 0
    for j, n in enumerate(sizes):
        au, av, aw = synth.rgg_graph(int(n), 6, 1, 255, 1000 + j)
        ro, col, w = synth.to_csr(int(n), au, av, aw, True)
        graphs.append(ctx.upload_graph(ro, col, w))
        ptrs.append(pool.offset(off))
        lds.append(ld_of[j])
        off += n * ld_of[j]
    ops = float(sum(2.0 * n ** 3 for n in sizes))

    def build():
        ctx.dense_from_graphs(graphs, ptrs, lds)

    def plain():
        for p_, n, ld in zip(ptrs, sizes, lds):
            ctx.fw(p_, int(n), int(ld), ctx.stream)

    tuned = _tune(ctx, args, "batch", fa.TuneCache.key_batch(sizes),
                  lambda: ctx.tune_batch(sizes, reps=1))
    plan = ctx.fw_plan(ptrs, [int(x) for x in sizes], lds, use_graph=True)
    info = plan.info()

    def batched():
        plan.run(ctx.stream)

    ns = [int(x) for x in sizes]
    pplan = ctx.fw_plan(ptrs, ns, lds, use_graph=True, persistent=True)

    def persistent():
        pplan.run(ctx.stream)

    res = {}
    runs = (("batched", batched),) if args.batched_only else (
        ("plain", plain), ("batched", batched), ("persistent", persistent))
    for name, fn in runs:
        build()
        fn()
        ctx.sync()
        t_ms = []
        for _ in range(args.batch_steps):
            build()
            ctx.sync()
            t0 = time.perf_counter()
            fn()
            ctx.sync()
            t_ms.append(1e3 * (time.perf_counter() - t0))
        res[name] = float(np.median(t_ms))
        if name == "persistent":
            pplan.check()   # no wait gave up in any run
        res[name + "_sum"] = ctx.checksum(pool.ptr, total)
        log(f"[bench] c4 {name}: {res[name]:.2f} ms")
    plan.close()
    pplan.close()
    for g in graphs:
        g.close()
    pool.free()
    same = None if args.batched_only else (res["plain_sum"] == res["batched_sum"] ==
                                           res["persistent_sum"])
    plain_ms = res.get("plain")
    pers_ms = res.get("persistent")
    bt = res["batched"]
    return {"workload": f"{args.jobs} random geometric graphs, n log-uniform in [64, 2048], "
                        f"deg 6, directed, w 1..255; full APSP per graph as one part (K=1: "
                        f"dense build + blocked FW, no boundary so no SSSP stage)",
            "jobs": int(args.jobs), "sum_n": int(sizes.sum()), "ops": ops,
            "plain_ms": None if plain_ms is None else round(plain_ms, 3),
            "batched_ms": round(bt, 3),
            "plain_GOp_s": None if plain_ms is None else round(ops / (plain_ms * 1e-3) / 1e9, 1),
            "batched_GOp_s": round(ops / (bt * 1e-3) / 1e9, 1),
            "batched_over_plain": None if plain_ms is None else round(plain_ms / bt, 2),
            "batched_frac_of_ceiling": round(ops / (bt * 1e-3) / 1e12 / ceiling, 3),
            "batched_frac_of_spec_peak": round(ops / (bt * 1e-3) / 1e12 / VALU_PEAK_TOPS, 3),
            "batched_launches": info["launches"], "batched_tasks": info["tasks"],
            "persistent_ms": None if pers_ms is None else round(pers_ms, 3),
            "persistent_GOp_s": None if pers_ms is None else round(ops / (pers_ms * 1e-3) / 1e9, 1),
            "persistent_over_plain": None if pers_ms is None else round(plain_ms / pers_ms, 2),
            "steps": args.batch_steps, "identical": same, "tuned": tuned}


def bench_batch(args):
    """c4 alone."""
    import fastapsp_amd as fa
    lib = fa.load()
    ctx = fa.Context(0)
    ceiling = ctx.probe_minplus_ceiling()
    out = batch_section(args, ctx, lib, ceiling)
    ctx.close()
    res = {"metric": "min-plus GOp/s", "value": out["batched_GOp_s"], "unit": "GOp/s",
           "n_gpus": 1, "steps": args.batch_steps, "warmup": 1,
           "ms_per_step": out["batched_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32", "data": "synthetic",
           "config": {"workload": out["workload"]}, "minplus_ceiling_TOPs": round(ceiling, 2),
           "batch_c4": out}
    print(json.dumps(res), flush=True)
    return 1 if out["identical"] is False else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--cpu-rows", type=int, default=16,
                    help="rows of the product timed for the CPU baseline (0: none)")
    ap.add_argument("--cpu-fw-n", type=int, default=1024,
                    help="leading block closed by the CPU FW for the baseline")
    ap.add_argument("--dense-only", action="store_true",
                    help="default workload: only the dense step (no c4/c5/c3 sections)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--grid-steps", type=int, default=1)
    ap.add_argument("--workload", default="dense",
                    choices=["dense", "copy", "c1", "grid", "gridbatch", "fwsweep",
                             "ssspsweep", "batch", "serve"])
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--batched-only", action="store_true",
                    help="batch workload: time the batched closure only")
    ap.add_argument("--sizes", default="256,512,1024,2048,4096,8192")
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--parts-k", type=int, default=256)
    ap.add_argument("--max-parts", type=int, default=0)
    ap.add_argument("--group", type=int, default=32)
    ap.add_argument("--call-parts", type=int, default=128)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=2500.0,
                    help="served workload: fixed Poisson arrival rate, requests per second")
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--hbm-gib", type=float, default=16.0)
    ap.add_argument("--query-parts", type=int, default=256,
                    help="parts of the served resident graph's query index (0: SSSP per query)")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: loaded if present, missing buckets tuned and saved")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the parts of a "
              "graph shard across GPUs without exchange", file=sys.stderr)
        return 2
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "grid":
        return bench_grid_pipeline(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "c1":
        return bench_c1(args)
    if args.workload == "gridbatch":
        return bench_grid_batch(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "fwsweep":
        return bench_fw_sweep(args)
    if args.workload == "ssspsweep":
        return bench_sssp_sweep(args)
    return bench_dense(args)


if __name__ == "__main__":
    sys.exit(main())
