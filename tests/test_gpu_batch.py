"""Batched closures (phase-synchronous launches over size-ordered task
queues, pivot rounds in pairs) and the
batched dense build, bit-exact against the oracle and against the one-matrix
path."""

import numpy as np
import pytest

import oracle_lib as O
from fastapsp_amd import synth

pytestmark = pytest.mark.gpu


def dense_of(V, ro, col, w):
    m = np.full((V, V), O.INF, np.int32)
    src = np.repeat(np.arange(V), np.diff(ro))
    for s_, c_, w_ in zip(src, col, w):   # CSR order: the last duplicate wins
        m[s_, c_] = w_
    m[np.arange(V), np.arange(V)] = 0
    return m


def test_fw_batched_uneven_sizes_and_strides(gpu_ctx):
    sizes = [1, 5, 63, 64, 65, 127, 200, 300, 513, 1000, 64, 2, 129]
    mats, devs, ptrs, lds = [], [], [], []
    for k, n in enumerate(sizes):
        m = synth.dense_matrix(n, min(1.0, 5.0 / n), 1, 1000, 500 + k)
        ld = n + (k % 3) * 3
        buf = np.full((n, ld), -7, np.int32)
        buf[:, :n] = m
        d = gpu_ctx.to_device(buf)
        mats.append(m)
        devs.append(d)
        ptrs.append(d.ptr)
        lds.append(ld)
    gpu_ctx.fw_batched(ptrs, sizes, lds)
    for n, m, d, ld in zip(sizes, mats, devs, lds):
        got = d.get().reshape(n, ld)
        np.testing.assert_array_equal(got[:, :n], O.floyd(m), err_msg=f"n={n}")
        assert (got[:, n:] == -7).all()
        d.free()


def test_dense_from_graphs_and_batched_apsp(gpu_ctx):
    sizes = synth.loguniform_sizes(24, 40, 600, 7)
    graphs, hosts, devs = [], [], []
    for j, n in enumerate(sizes):
        n = int(n)
        au, av, aw = synth.rgg_graph(n, 6, 1, 255, 900 + j)
        if j == 0 and len(au):
            # a duplicate arc (the later one wins) and a self loop (diagonal stays 0)
            au = np.concatenate([au, au[:1], [0]])
            av = np.concatenate([av, av[:1], [0]])
            aw = np.concatenate([aw, [7], [3]])
        ro, col, w = synth.to_csr(n, au, av, aw, True)
        graphs.append(gpu_ctx.upload_graph(ro, col, w))
        hosts.append((n, ro, col, w))
        devs.append(gpu_ctx.alloc((n, n)))
    gpu_ctx.dense_from_graphs(graphs, [d.ptr for d in devs], [h[0] for h in hosts])
    for (n, ro, col, w), d in zip(hosts, devs):
        np.testing.assert_array_equal(d.get(), dense_of(n, ro, col, w))
    gpu_ctx.fw_batched([d.ptr for d in devs], [h[0] for h in hosts], [h[0] for h in hosts])
    for (n, ro, col, w), d in zip(hosts, devs):
        np.testing.assert_array_equal(d.get(), O.floyd(dense_of(n, ro, col, w)))
    for g in graphs:
        g.close()


def _plan_launches(sizes):
    """Launches of a plan: 8+ jobs split (size order, alternately) into two
    chains; a chain has as many launches as its longest job has phases
    (pivot rounds in pairs of 7 phases -- diag, panels, cross, diag,
    panels, 128-deep rest, cross -- and 3 for an odd last round)."""
    def phases(n):
        nb = (n + 63) // 64
        return 1 if nb == 1 else 7 * (nb // 2) + 3 * (nb % 2)
    order = sorted(sizes, reverse=True)
    chains = 2 if len(sizes) >= 8 else 1
    return sum(max(phases(n) for n in order[c::chains]) for c in range(chains))


@pytest.mark.parametrize("sizes", [[300, 64, 1, 129, 700, 65],
                                   [300, 64, 1, 129, 700, 65, 200, 513, 90, 1000]],
                         ids=["one-chain", "two-chains"])
def test_fw_plan_graph_replays_match_eager(gpu_ctx, sizes):
    """The captured phase sequence (one stream, or two chains of jobs on two
    streams joined in the graph), replayed several times with eager work and
    synchronisation in between, gives the eager result every time."""
    import fastapsp_amd as fa
    lib = fa.load()
    srcs, work = [], []
    for k, n in enumerate(sizes):
        m = synth.dense_matrix(n, min(1.0, 5.0 / n), 1, 1000, 700 + k)
        srcs.append((m, gpu_ctx.to_device(m)))
        work.append(gpu_ctx.alloc((n, n)))
    ptrs, lds = [w.ptr for w in work], list(sizes)

    def restore():
        for (m, s), w, n in zip(srcs, work, sizes):
            fa.check(lib.fap_copy_i32(gpu_ctx.handle, w.ptr, s.ptr, n * n, None))

    restore()
    gpu_ctx.fw_batched(ptrs, sizes, lds)
    eager = [w.get() for w in work]
    for (m, _), e in zip(srcs, eager):
        np.testing.assert_array_equal(e, O.floyd(m))
    plan = gpu_ctx.fw_plan(ptrs, sizes, lds, use_graph=True)
    info = plan.info()
    assert info["launches"] == _plan_launches(sizes)
    plan.check()   # no run yet
    for rep in range(3):
        restore()
        gpu_ctx.sync()
        plan.run()
        plan.check()   # a phase-launch plan: waits for the run, nothing to report
        for w, e in zip(work, eager):
            np.testing.assert_array_equal(w.get(), e, err_msg=f"replay {rep}")
    plan.close()


def test_fw_batched_paired_rounds_match_single_graph_fw(gpu_ctx):
    """Many pivot-round pairs (block counts 33, 32, 17, 16, 3, 3, 2, 1) and
    odd last rounds, ragged last blocks, padded strides: the batched closure
    equals the single-graph blocked FW (itself checked against the oracle)
    on every matrix."""
    sizes = [2049, 1985, 1088, 1023, 130, 192, 65, 64]
    work, plain, lds = [], [], []
    for k, n in enumerate(sizes):
        m = synth.dense_matrix(n, min(1.0, 6.0 / n), 1, 1000, 900 + k)
        ld = n + (4 if k % 2 else 0)
        buf = np.full((n, ld), -7, np.int32)
        buf[:, :n] = m
        work.append(gpu_ctx.to_device(buf))
        plain.append(gpu_ctx.to_device(buf))
        lds.append(ld)
    gpu_ctx.fw_batched([w.ptr for w in work], sizes, lds)
    for p, n, ld in zip(plain, sizes, lds):
        gpu_ctx.fw(p.ptr, n, ld, gpu_ctx.stream)
    gpu_ctx.sync()
    for n, w, p, ld in zip(sizes, work, plain, lds):
        got, want = w.get().reshape(n, ld), p.get().reshape(n, ld)
        np.testing.assert_array_equal(got, want, err_msg=f"n={n}")
        assert (got[:, n:] == -7).all()
        w.free()
        p.free()


def test_fw_plan_two_chains_with_an_empty_chain(gpu_ctx):
    """Eight jobs of which seven are empty: the second chain has no launch,
    and its branch of the captured graph is only the fork and the join;
    every replay still closes the one matrix exactly."""
    import fastapsp_amd as fa
    lib = fa.load()
    n = 200
    m = synth.dense_matrix(n, 5.0 / n, 1, 1000, 4242)
    src = gpu_ctx.to_device(m)
    work = gpu_ctx.alloc((n, n))
    sizes = [n] + [0] * 7
    ptrs = [work.ptr] + [0] * 7
    plan = gpu_ctx.fw_plan(ptrs, sizes, sizes, use_graph=True)
    assert plan.info()["launches"] == _plan_launches(sizes)
    want = O.floyd(m)
    for rep in range(3):
        fa.check(lib.fap_copy_i32(gpu_ctx.handle, work.ptr, src.ptr, n * n, None))
        gpu_ctx.sync()
        plan.run()
        gpu_ctx.sync()
        np.testing.assert_array_equal(work.get().reshape(n, n), want, err_msg=f"replay {rep}")
    plan.close()
    src.free()
    work.free()
