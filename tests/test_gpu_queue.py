"""The batched closure as ONE persistent launch (fap_fw_batched_persistent):
every phase of every matrix in one queue, tasks claimed in queue order,
in-kernel bounded waits on per-matrix completion counters.  Bit-exact
against the oracle and against the phase-synchronous launches
(fap_fw_batched); the queue's ordering and progress argument is checked on
the CPU by tests/cpp/batch_schedule_check.cpp."""

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import FapError, synth

pytestmark = pytest.mark.gpu

FAP_E_ARG, FAP_E_DEVICE = -1, -6


def _ld32(n):
    return max(32, (n + 31) // 32 * 32)


def _row(ctx, ptr, ld, r, n):
    out = np.empty(ld, np.int32)
    fa.check(fa.load().fap_memcpy_d2h(ctx.handle, out.ctypes.data, int(ptr) + 4 * r * ld,
                                      out.nbytes), "d2h row")
    return out[:n]


def _matrices(ctx, sizes, seed):
    mats, devs = [], []
    for k, n in enumerate(sizes):
        m = synth.dense_matrix(n, min(1.0, 5.0 / n), 1, 1000, seed + k)
        ld = _ld32(n)
        buf = np.full((n, ld), -7, np.int32)
        buf[:, :n] = m
        d = ctx.to_device(buf)
        assert d.ptr % 128 == 0
        mats.append(m)
        devs.append(d)
    return mats, devs


def test_persistent_uneven_sizes_match_oracle(gpu_ctx):
    sizes = [1, 5, 63, 64, 65, 127, 200, 300, 513, 1000, 64, 2, 129, 385]
    print(f"[queue] sizes {sizes}", flush=True)
    mats, devs = _matrices(gpu_ctx, sizes, 700)
    gpu_ctx.fw_batched_persistent([d.ptr for d in devs], sizes, [_ld32(n) for n in sizes])
    for n, m, d in zip(sizes, mats, devs):
        got = d.get().reshape(n, _ld32(n))
        np.testing.assert_array_equal(got[:, :n], O.floyd(m), err_msg=f"n={n}")
        assert (got[:, n:] == -7).all(), f"padding of n={n} written"
        d.free()


def test_persistent_matches_phase_launches_on_512_rggs(gpu_ctx):
    """The 512-graph uneven batch (random geometric graphs, n log-uniform in
    [64, 2048], degree 6, w 1..255) closed three ways -- per graph, by the
    phase-synchronous launches and by the single launch, twice -- with
    identical checksums; sampled rows against oracle Dijkstra."""
    ctx = gpu_ctx
    sizes = synth.loguniform_sizes(512, 64, 2048, 4).astype(int)
    lds = [_ld32(int(n)) for n in sizes]
    total = sum(int(n) * ld for n, ld in zip(sizes, lds))
    pool = ctx.alloc(total)
    assert pool.ptr % 128 == 0
    graphs, csrs, ptrs, off = [], [], [], 0
    for j, n in enumerate(sizes):
        au, av, aw = synth.rgg_graph(int(n), 6, 1, 255, 1000 + j)
        csr = synth.to_csr(int(n), au, av, aw, True)
        csrs.append(csr)
        graphs.append(ctx.upload_graph(*csr))
        ptrs.append(pool.offset(off))
        off += int(n) * lds[j]
    ns = [int(x) for x in sizes]
    print(f"[queue] 512 graphs, sum n = {sizes.sum()}", flush=True)
    ctx.dense_from_graphs(graphs, ptrs, lds)
    for p_, n, ld in zip(ptrs, ns, lds):
        ctx.fw(p_, n, ld)
    ctx.sync()
    want = ctx.checksum(pool.ptr, total)
    ctx.dense_from_graphs(graphs, ptrs, lds)
    ctx.sync()
    ctx.fw_batched(ptrs, ns, lds)
    assert ctx.checksum(pool.ptr, total) == want, "phase launches"
    for rep in range(2):
        ctx.dense_from_graphs(graphs, ptrs, lds)
        ctx.sync()
        ctx.fw_batched_persistent(ptrs, ns, lds)
        assert ctx.checksum(pool.ptr, total) == want, f"single launch, run {rep}"
    plan = ctx.fw_plan(ptrs, ns, lds, use_graph=True, persistent=True)
    try:
        assert plan.info()["launches"] == 1
        for rep in range(2):   # graph replays, eager work in between
            ctx.dense_from_graphs(graphs, ptrs, lds)
            ctx.sync()
            plan.run(ctx.stream)
            plan.check()
            assert ctx.checksum(pool.ptr, total) == want, f"persistent plan, replay {rep}"
    finally:
        plan.close()
    rng = np.random.default_rng(45)
    for j in rng.choice(len(ns), 12, replace=False):
        n = ns[j]
        ro, col, w = csrs[j]
        for r in np.unique(np.r_[0, n - 1, rng.integers(0, n, 1)]):
            got = _row(ctx, ptrs[j], lds[j], int(r), n)
            np.testing.assert_array_equal(got, O.batched_sssp(np.array([r], np.int32), ro, col, w)
                                          .reshape(-1), err_msg=f"graph {j} row {r}")
    for g in graphs:
        g.close()
    pool.free()


def test_persistent_refuses_rows_not_line_aligned(gpu_ctx):
    d = gpu_ctx.alloc((70, 72))
    with pytest.raises(FapError) as e:
        gpu_ctx.fw_batched_persistent([d.ptr], [70], [72])   # ld not a multiple of 32
    assert e.value.code == FAP_E_ARG and "128-byte" in str(e.value)
    with pytest.raises(FapError) as e:
        gpu_ctx.fw_batched_persistent([d.ptr + 64], [40], [64])  # base not 128-byte aligned
    assert e.value.code == FAP_E_ARG
    d.free()


def test_persistent_wait_limit_gives_up_and_context_recovers(gpu_ctx):
    """A 1-us wait limit (forced from the host through the option): the first
    task that must wait for its matrix's earlier phases gives up, every
    workgroup exits, the call reports FAP_E_DEVICE; with the default limit
    the same batch then closes exactly."""
    sizes = [1024, 300, 64]
    mats, devs = _matrices(gpu_ctx, sizes, 810)
    ptrs, lds = [d.ptr for d in devs], [_ld32(n) for n in sizes]
    gpu_ctx.set_option("batch_wait_us", 1)
    try:
        with pytest.raises(FapError) as e:
            gpu_ctx.fw_batched_persistent(ptrs, sizes, lds)
        assert e.value.code == FAP_E_DEVICE
    finally:
        gpu_ctx.set_option("batch_wait_us", 0)
    for m, d, n in zip(mats, devs, sizes):
        buf = np.full((n, _ld32(n)), -7, np.int32)
        buf[:, :n] = m
        d.set(buf)
    gpu_ctx.fw_batched_persistent(ptrs, sizes, lds)
    for n, m, d in zip(sizes, mats, devs):
        got = d.get().reshape(n, _ld32(n))
        np.testing.assert_array_equal(got[:, :n], O.floyd(m), err_msg=f"n={n}")
        d.free()


@pytest.mark.parametrize("sizes", [[0, 70, 0, 1], [640], [5, 17, 33, 64, 2, 9, 40, 63, 11, 64, 1]])
def test_persistent_edge_batches(gpu_ctx, sizes):
    """Empty matrices (null pointers allowed), one matrix (one queue: the
    workgroups of the other XCDs exit at once), and more matrices than XCDs
    that are all single-block (every task a diagonal with nothing to wait
    for)."""
    print(f"[queue] sizes {sizes}", flush=True)
    live = [n for n in sizes if n > 0]
    mats, devs = _matrices(gpu_ctx, live, 900)
    it = iter(devs)
    ptrs = [next(it).ptr if n > 0 else 0 for n in sizes]
    gpu_ctx.fw_batched_persistent(ptrs, sizes, [_ld32(n) if n > 0 else 0 for n in sizes])
    for n, m, d in zip(live, mats, devs):
        got = d.get().reshape(n, _ld32(n))
        np.testing.assert_array_equal(got[:, :n], O.floyd(m), err_msg=f"n={n}")
        assert (got[:, n:] == -7).all()
        d.free()


@pytest.mark.parametrize("use_graph", [True, False])
def test_persistent_plan_replays_match_oracle(gpu_ctx, use_graph):
    """A persistent plan (one launch per run, queues and task map uploaded
    once), replayed three times on reinitialised matrices, with and without
    its captured graph; check() after each run."""
    sizes = [300, 64, 1, 129, 700, 65, 200, 33, 513]
    print(f"[queue] plan sizes {sizes} graph={use_graph}", flush=True)
    mats, devs = _matrices(gpu_ctx, sizes, 950)
    ptrs, lds = [d.ptr for d in devs], [_ld32(n) for n in sizes]
    want = [O.floyd(m) for m in mats]
    plan = gpu_ctx.fw_plan(ptrs, sizes, lds, use_graph=use_graph, persistent=True)
    try:
        plan.check()   # no run yet: nothing to report
        info = plan.info()
        assert info["launches"] == 1 and info["ops"] == sum(2 * n ** 3 for n in sizes)
        for rep in range(3):
            for m, d, n in zip(mats, devs, sizes):
                buf = np.full((n, _ld32(n)), -7, np.int32)
                buf[:, :n] = m
                d.set(buf)
            plan.run(gpu_ctx.stream)
            plan.check()
            for n, w, d in zip(sizes, want, devs):
                got = d.get().reshape(n, _ld32(n))
                np.testing.assert_array_equal(got[:, :n], w, err_msg=f"replay {rep} n={n}")
                assert (got[:, n:] == -7).all()
    finally:
        plan.close()
    for d in devs:
        d.free()
