"""The uneven batch, the served stream and the full grid pipeline at their own
full sizes on the GPU, checked against the oracle (Dijkstra rows, pinned on
scipy and the reference's own dijkstra_path) -- bounded by sampling rows,
not by shrinking the workloads.

* 512 random geometric graphs, n log-uniform in [64, 2048], average degree 6,
  directed, w 1..255: each closed to full APSP by the batched closure plan
  replayed from a hipGraph (the per-graph pipeline at K = 1,
  fap/fap.cpp:92-181, with no boundary vertices).
* A Poisson stream at a fixed rate of whole-graph requests (RGG, n
  log-uniform in [256, 4096]) and 1-64-pair query batches against the
  resident 512x512 grid, through the batching server with the query index
  over 256 parts.
* The 512x512 grid (10% of edges dropped, 2% diagonal shortcuts, w 1..100),
  K = 256: every part through fapGraph::solveSubGraphs in its default
  pipelined form and in the plain part-by-part form (fap/fap.cpp:197-207),
  checksums equal part by part, sampled rows equal to Dijkstra.
"""

import threading
import time

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _rows(ctx, ptr, ld, r0, nrows, n):
    ptr, ld, r0, nrows = int(ptr), int(ld), int(r0), int(nrows)
    out = np.empty((nrows, ld), np.int32)
    fa.check(fa.load().fap_memcpy_d2h(ctx.handle, out.ctypes.data, ptr + 4 * r0 * ld,
                                      out.nbytes), "d2h rows")
    return out[:, :n]


def test_uneven_batch_512_rggs_full_workload(gpu_ctx):
    ctx = gpu_ctx
    sizes = synth.loguniform_sizes(512, 64, 2048, 4).astype(int)
    assert sizes.min() >= 64 and sizes.max() <= 2048
    csrs, graphs, ptrs, lds = [], [], [], []
    ld_of = [(n + 3) // 4 * 4 for n in sizes]
    total = sum(n * ld for n, ld in zip(sizes, ld_of))
    pool = ctx.alloc(total)
    off = 0
    for j, n in enumerate(sizes):
        au, av, aw = synth.rgg_graph(int(n), 6, 1, 255, 1000 + j)
        csr = synth.to_csr(int(n), au, av, aw, True)
        csrs.append(csr)
        graphs.append(ctx.upload_graph(*csr))
        ptrs.append(pool.offset(off))
        lds.append(int(ld_of[j]))
        off += int(n) * int(ld_of[j])
    print(f"[c4] 512 graphs, sum n = {sizes.sum()}, {4 * total / 2**30:.2f} GiB", flush=True)
    # eager: every graph closed on its own (the plain loop)
    ctx.dense_from_graphs(graphs, ptrs, lds)
    for p_, n, ld in zip(ptrs, sizes, lds):
        ctx.fw(p_, int(n), int(ld))
    ctx.sync()
    want = ctx.checksum(pool.ptr, total)
    plan = ctx.fw_plan(ptrs, [int(x) for x in sizes], lds, use_graph=True)
    try:
        for rep in range(3):   # replays, with eager work in between
            ctx.dense_from_graphs(graphs, ptrs, lds)
            ctx.sync()
            plan.run(ctx.stream)
            ctx.sync()
            assert ctx.checksum(pool.ptr, total) == want, f"replay {rep}"
    finally:
        plan.close()
    rng = np.random.default_rng(44)
    full = 0
    for j, n in enumerate(sizes):
        n = int(n)
        ro, col, w = csrs[j]
        if n < 300:
            got = _rows(ctx, ptrs[j], lds[j], 0, n, n)
            np.testing.assert_array_equal(got, O.batched_sssp(np.arange(n, dtype=np.int32),
                                                              ro, col, w), err_msg=f"graph {j}")
            full += 1
        else:
            for r in np.unique(np.r_[0, n - 1, rng.integers(0, n, 2)]):
                got = _rows(ctx, ptrs[j], lds[j], int(r), 1, n)
                np.testing.assert_array_equal(got, O.batched_sssp(np.array([r], np.int32),
                                                                  ro, col, w),
                                              err_msg=f"graph {j} row {r}")
    assert full > 150
    for g in graphs:
        g.close()
    pool.free()


def _grid512():
    u, v, w = synth.grid_graph(512, 512, 0.10, 0.02, 1, 100, 3)
    return synth.to_csr(512 * 512, u, v, w, False)


def test_served_poisson_mix_against_resident_grid():
    ro, col, w = _grid512()
    V = len(ro) - 1
    rng = np.random.default_rng(5)
    count = 96
    kinds = rng.random(count) < 0.5
    sizes = synth.loguniform_sizes(count, 256, 4096, 55)
    pool = []
    for i in range(count):
        if kinds[i]:
            n = int(sizes[i])
            au, av, aw = synth.rgg_graph(n, 6, 1, 255, 7000 + i)
            pool.append(("graph", synth.to_csr(n, au, av, aw, True)))
        else:
            k = int(rng.integers(1, 65))
            pool.append(("queries", (rng.integers(0, V, k).astype(np.int32),
                                     rng.integers(0, V, k).astype(np.int32))))
    rate = 2500.0
    gaps = np.random.default_rng(11).exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    ids = [None] * count
    ready = threading.Semaphore(0)
    got = []
    with fa.Server(hbm_budget=8 << 30, lanes=3, query_parts=256) as srv:
        srv.attach_graph(ro, col, w)
        print(f"[c5] attached V={V}, index {srv.stats()['index_bytes'] / 2**30:.1f} GiB",
              flush=True)

        def submitter():
            t0 = time.perf_counter()
            for i, (kind, payload) in enumerate(pool):
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[i] = (srv.submit_graph(*payload) if kind == "graph"
                          else srv.submit_queries(*payload))
                ready.release()

        th = threading.Thread(target=submitter)
        th.start()
        for i, (kind, payload) in enumerate(pool):
            ready.acquire()
            info = srv.wait(ids[i])
            assert info["state"] == "done", (i, info)
            if kind == "graph":
                n = len(payload[0]) - 1
                rs = np.unique(np.r_[0, n - 1, rng.integers(0, n, 6)]).astype(np.int32)
                got.append((info, [(int(r), srv.rows(ids[i], int(r), 1)[0]) for r in rs]))
            else:
                got.append((info, srv.answers(ids[i])))
            srv.release(ids[i])
        th.join()
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == count
    assert max(g[0]["wave_size"] for g in got) >= 2          # waves did batch requests
    # oracle: graph rows, then every query pair through Dijkstra rows of its sources
    for (kind, payload), (info, res) in zip(pool, got):
        if kind == "graph":
            g_ro, g_col, g_w = payload
            srcs = np.array([r for r, _ in res], np.int32)
            want = O.batched_sssp(srcs, g_ro, g_col, g_w)
            for k, (r, row) in enumerate(res):
                np.testing.assert_array_equal(row, want[k])
    qsrc = np.unique(np.concatenate([p[1][0] for p in pool if p[0] == "queries"]))
    rows = O.batched_sssp(qsrc.astype(np.int32), ro, col, w)
    where = {int(s): k for k, s in enumerate(qsrc)}
    for (kind, payload), (info, res) in zip(pool, got):
        if kind == "queries":
            src, dst = payload
            want = np.array([rows[where[int(s)], int(t)] for s, t in zip(src, dst)], np.int32)
            np.testing.assert_array_equal(res, want)


def test_grid512_all_parts_pipelined_equals_plain_and_dijkstra():
    ro, col, w = _grid512()
    V = len(ro) - 1
    K = 256
    rng = np.random.default_rng(9)
    checked = []
    with fa.Graph.from_csr(ro, col, w.astype(np.float32), K) as g:
        g.preCondition()
        assert g.isSplit()
        vn, vo, bn, _ = g.part_info()
        st2ed = g.getMapping()
        for c0 in range(1, K + 1, 64):
            chunk = list(range(c0, min(c0 + 64, K + 1)))
            t0 = time.time()
            g.setBatched(True)              # the default: pipelined groups
            g.solveSubGraphs(chunk)
            piped = {p: g.checksum(p) for p in chunk}
            sample = {}
            for p in chunk:
                n = int(vn[p])
                rs = np.unique(np.r_[0, max(int(bn[p]) - 1, 0), n - 1, rng.integers(0, n)])
                sample[p] = [(int(r), g.getSubGraphRowsInt32(p, int(r), 1)[0]) for r in rs]
            g.setBatched(False)             # plain: one part after another
            g.solveSubGraphs(chunk)
            for p in chunk:
                assert g.checksum(p) == piped[p], f"part {p}"
            print(f"[c3] parts {chunk[0]}..{chunk[-1]}: {time.time() - t0:.1f} s", flush=True)
            for p in chunk:
                for r, row in sample[p]:
                    checked.append((int(st2ed[int(vo[p]) + r]), row))
    srcs = np.array([s for s, _ in checked], np.int32)
    want = O.batched_sssp(srcs, ro, col, w)
    for k, (s, row) in enumerate(checked):
        np.testing.assert_array_equal(row, want[k], err_msg=f"source {s}")
    assert len(checked) >= 3 * K
