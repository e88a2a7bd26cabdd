#!/usr/bin/env python3
"""Benchmark of the subgraph-solve hot path on one MI355X.

Default workload (the configuration the headline metric is quoted on): the
dense kernel pair on an n = 8192 int32 distance matrix -- each off-diagonal
entry an arc with probability 0.01 and weight in [1, 1000], else INF,
diagonal 0 -- i.e. one blocked Floyd-Warshall closure of the matrix in place
plus one min-plus product of two such matrices.  One step = restore the FW
input (device copy) + FW + min-plus, inputs resident in HBM.
value = min-plus GOp/s = (2 n^3 [FW] + 2 n^3 [product]) / step time, 2 ops
(add, min) per update as the reference counts (fap/kernel/cuda/floyd.cu:340,
fap/kernel/cuda/minplus.cu:483).

The roofline object is for the min-plus product kernel, timed with HIP events
on its own stream over the timed steps; the Floyd-Warshall closure is
reported beside it.  Peak = VALU int32: 256 CUs x 64 lanes x 2.4 GHz =
39.3 T int32 ops/s (min-plus is not a multiply-accumulate, so no MFMA).

cpu_baseline: the reference's own CPU code (oracle/_ref, compiled from its
headers; min_plus_path_CPU_advanced, single-threaded like the reference) on
a row band of the same product, scaled to GOp/s; the int32 restatement
(oracle/_build) is used where the reference build is absent.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--n 8192]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.32 T int32 ops/s
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline_minplus(n: int, A: np.ndarray, B: np.ndarray, rows: int):
    """GOp/s of the reference CPU min-plus on `rows` rows of the n^3 product."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    a = np.ascontiguousarray(A[:rows])
    if O.ref_available():
        fa = O.to_ref(a)
        fb = O.to_ref(B)
        Cm = np.empty((rows, n), np.float32)
        bp = np.zeros((n, n), np.int32)
        cp = np.empty((rows, n), np.int32)
        fn = O.ref().ref_min_plus_cpu_advanced
        t = time.perf_counter()
        fn(O._f(fa), O._f(fb), O._p(bp), O._f(Cm), O._p(cp), C.c_int32(rows), C.c_int32(n),
           C.c_int32(n))
        dt = time.perf_counter() - t
        kind = "reference"
    else:
        t = time.perf_counter()
        O.min_plus(a, B)
        dt = time.perf_counter() - t
        kind = "port"
    ops = 2.0 * rows * n * n
    return {"value": ops / dt / 1e9, "unit": "GOp/s", "cores": 1, "kind": kind,
            "sample": f"min_plus_path_CPU_advanced on {rows} of {n} rows of the {n}^3 product "
                      f"({ops:.3g} ops, {dt:.1f} s), scaled to GOp/s"}


def bench_dense(args):
    import fastapsp_amd as fa
    from fastapsp_amd import synth
    lib = fa.load()
    n = args.n
    log(f"[bench] generating n={n} matrices")
    t0 = time.time()
    M0 = synth.dense_matrix(n, 0.01, 1, 1000, 2)
    A = synth.dense_matrix(n, 0.01, 1, 1000, 21)
    B = synth.dense_matrix(n, 0.01, 1, 1000, 22)
    log(f"[bench] generated in {time.time() - t0:.1f}s")
    ctx = fa.Context(0)
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
    ops_step = 4.0 * n ** 3
    ms_step = 1e3 * total / args.steps
    value = ops_step / (total / args.steps) / 1e9
    mp_avg = float(np.mean(mp_ms))
    fw_avg = float(np.mean(fw_ms))
    ach = 2.0 * n ** 3 / (mp_avg * 1e-3) / 1e12
    ach_fw = 2.0 * n ** 3 / (fw_avg * 1e-3) / 1e12
    cpu = None
    if args.cpu_rows > 0:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline_minplus(n, A, B, rows=args.cpu_rows)
    res = {
        "metric": "min-plus GOp/s",
        "value": round(value, 1),
        "unit": "GOp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"dense FW + min-plus product, n={n}, arc p=0.01, w 1..1000",
                   "n": n, "ops_per_step": ops_step},
        "roofline": {"bound": "valu", "kernel": "minplus_gemm_kernel",
                     "achieved": round(ach, 2), "peak": round(VALU_PEAK_TOPS, 2),
                     "unit": "TOP/s", "frac": round(ach / VALU_PEAK_TOPS, 3), "traffic": None,
                     "ms_per_launch": round(mp_avg, 3)},
        "roofline_fw": {"bound": "valu", "kernels": "fw_diag/fw_panel/fw_rest",
                        "achieved": round(ach_fw, 2), "peak": round(VALU_PEAK_TOPS, 2),
                        "unit": "TOP/s", "frac": round(ach_fw / VALU_PEAK_TOPS, 3),
                        "ms_per_closure": round(fw_avg, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--cpu-rows", type=int, default=24)
    ap.add_argument("--workload", default="dense", choices=["dense", "copy"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the parts of a "
              "graph shard across GPUs without exchange, see DESIGN.md", file=sys.stderr)
        return 2
    if args.workload == "copy":
        return bench_copy(args)
    return bench_dense(args)


if __name__ == "__main__":
    sys.exit(main())
