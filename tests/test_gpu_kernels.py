"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

Integer work: every comparison is bit-exact.  Shapes cover ragged edges
(M, N, K not multiples of the 128/16 tiles), K = 0, INF-heavy operands,
accumulate mode, non-minimal and unaligned leading dimensions.
"""

import numpy as np
import pytest

import oracle_lib as O
from fastapsp_amd import synth
import fastapsp_amd as fa

pytestmark = pytest.mark.gpu
INF = O.INF


def rand_mat(rows, cols, seed, lo=0, hi=1000, p_inf=0.2):
    r = synth.uniform_int(synth.stream(seed, 1, rows * cols), lo, hi).reshape(rows, cols)
    m = synth.uniform01(synth.stream(seed, 2, rows * cols)).reshape(rows, cols) < p_inf
    r[m] = INF
    return r.astype(np.int32)


def padded(ctx, host, ld):
    """Device copy of host (rows x cols) with row stride ld >= cols."""
    rows, cols = host.shape
    buf = np.full((rows, ld), -7, np.int32)
    buf[:, :cols] = host
    return ctx.to_device(buf)


def unpad(dev, rows, cols):
    return dev.get().reshape(rows, -1)[:, :cols]


# ---------------- min-plus ----------------

MP_SHAPES = [(1, 1, 1), (5, 7, 3), (128, 128, 16), (129, 257, 17), (64, 300, 2), (300, 1000, 124),
             (257, 129, 250), (33, 4099, 61), (200, 300, 115), (70, 140, 15), (9, 600, 31),
             (909, 1100, 115), (192, 200, 40),   # 64-row tails: 13 rows, exactly 64 rows
             (16, 300, 40), (144, 1000, 7), (130, 257, 70), (1, 5000, 200)]  # up to 16 rows


@pytest.mark.parametrize("M,N,K", MP_SHAPES)
def test_minplus_matches_oracle(gpu_ctx, M, N, K):
    A = rand_mat(M, K, M * 7 + K)
    B = rand_mat(K, N, N * 3 + K)
    dA, dB = gpu_ctx.to_device(A), gpu_ctx.to_device(B)
    dC = gpu_ctx.alloc((M, N))
    gpu_ctx.minplus(dA.ptr, K, dB.ptr, N, dC.ptr, N, M, N, K)
    gpu_ctx.sync()
    np.testing.assert_array_equal(dC.get(), O.min_plus(A, B))
    for d in (dA, dB, dC):
        d.free()


def test_minplus_golden(gpu_ctx):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "minplus_13x37x29.npz"))
    A, B = g["A"], g["B"]
    dA, dB = gpu_ctx.to_device(A), gpu_ctx.to_device(B)
    dC = gpu_ctx.alloc((13, 29))
    gpu_ctx.minplus(dA.ptr, 37, dB.ptr, 29, dC.ptr, 29, 13, 29, 37)
    np.testing.assert_array_equal(dC.get(), g["C"])


@pytest.mark.parametrize("M,N,K,lda,ldb,ldc", [(70, 90, 33, 35, 93, 91), (128, 256, 32, 36, 260, 264),
                                                (17, 130, 5, 5, 131, 133),
                                                (150, 90, 33, 35, 93, 91),
                                                (12, 300, 130, 131, 301, 303),
                                                (141, 90, 70, 71, 93, 91)])
def test_minplus_strided_accumulate(gpu_ctx, M, N, K, lda, ldb, ldc):
    A = rand_mat(M, K, 101)
    B = rand_mat(K, N, 102)
    C0 = rand_mat(M, N, 103, hi=3000, p_inf=0.3)
    dA, dB, dC = padded(gpu_ctx, A, lda), padded(gpu_ctx, B, ldb), padded(gpu_ctx, C0, ldc)
    gpu_ctx.minplus(dA.ptr, lda, dB.ptr, ldb, dC.ptr, ldc, M, N, K, accumulate=True)
    got = dC.get().reshape(M, ldc)
    np.testing.assert_array_equal(got[:, :N], O.min_plus_acc(A, B, C0))
    assert (got[:, N:] == -7).all(), "wrote outside the row"


def test_minplus_unaligned_pointers(gpu_ctx):
    M, N, K = 50, 77, 19
    A = rand_mat(M, K, 201)
    B = rand_mat(K, N, 202)
    dA = gpu_ctx.to_device(np.concatenate([[0], A.ravel()]))
    dB = gpu_ctx.to_device(np.concatenate([[0, 0, 0], B.ravel()]))
    dC = gpu_ctx.alloc(M * N + 1)
    gpu_ctx.minplus(dA.offset(1), K, dB.offset(3), N, dC.offset(1), N, M, N, K)
    np.testing.assert_array_equal(dC.get()[1:].reshape(M, N), O.min_plus(A, B))


def test_minplus_k0_and_all_inf(gpu_ctx):
    M, N = 40, 50
    dC = gpu_ctx.to_device(np.zeros((M, N), np.int32))
    gpu_ctx.minplus(0, 1, 0, N, dC.ptr, N, M, N, 0)
    assert (dC.get() == INF).all()
    A = np.full((M, 8), INF, np.int32)
    B = np.full((8, N), INF, np.int32)
    dA, dB = gpu_ctx.to_device(A), gpu_ctx.to_device(B)
    gpu_ctx.minplus(dA.ptr, 8, dB.ptr, N, dC.ptr, N, M, N, 8)
    assert (dC.get() == INF).all()


def test_minplus_rejects_bad_arguments(gpu_ctx):
    import fastapsp_amd as fa
    dC = gpu_ctx.alloc((4, 4))
    with pytest.raises(fa.FapError):
        gpu_ctx.minplus(dC.ptr, 4, dC.ptr, 4, dC.ptr, 3, 4, 4, 4)  # ldc < N
    with pytest.raises(fa.FapError):
        gpu_ctx.minplus(dC.ptr, 4, dC.ptr, 4, dC.ptr, 4, -1, 4, 4)


# ---------------- Floyd-Warshall ----------------

@pytest.fixture
def fw_config(gpu_ctx, request):
    tile, look = request.param
    gpu_ctx.set_option("fw_tile", tile)
    gpu_ctx.set_option("fw_lookahead", look)
    yield request.param
    gpu_ctx.set_option("fw_tile", 0)
    gpu_ctx.set_option("fw_lookahead", -1)


FW_CONFIGS = [(64, 0), (128, 0), (64, 1), (128, 1)]


@pytest.mark.parametrize("fw_config", FW_CONFIGS, indirect=True, ids=lambda c: f"T{c[0]}-la{c[1]}")
@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 127, 128, 129, 200, 300, 513])
def test_fw_matches_oracle(gpu_ctx, fw_config, n):
    m = synth.dense_matrix(n, min(1.0, 6.0 / n), 1, 1000, n)
    d = gpu_ctx.to_device(m)
    gpu_ctx.fw(d.ptr, n, n)
    np.testing.assert_array_equal(d.get(), O.floyd(m))


def test_fw_fused_diagonal_matches_unfused(gpu_ctx):
    """n = 3100 (49 blocks of 64, the last one 28 wide): big enough that each
    round's remaining-tile launch also closes the next diagonal block (at
    least 8 tiles per CU).  The whole closure equals the unfused forms
    (look-ahead on, and 128-wide blocks), and sampled rows equal Dijkstra
    rows of the matrix graph."""
    n = 3100
    m = synth.dense_matrix(n, 0.004, 1, 1000, 31)
    ref = None
    for tile, look in ((64, 0), (64, 1), (128, 0)):
        gpu_ctx.set_option("fw_tile", tile)
        gpu_ctx.set_option("fw_lookahead", look)
        try:
            d = gpu_ctx.to_device(m)
            gpu_ctx.fw(d.ptr, n, n)
            got = d.get()
            d.free()
        finally:
            gpu_ctx.set_option("fw_tile", 0)
            gpu_ctx.set_option("fw_lookahead", -1)
        if ref is None:
            ref = got
        else:
            np.testing.assert_array_equal(got, ref)
    rows = np.array([0, 1500, 3071, 3099])
    mask = (m < O.INF) & ~np.eye(n, dtype=bool)
    src, dst = np.nonzero(mask)
    ro = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(src, minlength=n), out=ro[1:])
    want = O.batched_sssp(rows, ro, dst.astype(np.int32), m[src, dst].astype(np.int32))
    np.testing.assert_array_equal(ref[rows], want)


@pytest.mark.parametrize("n", [4200, 4500])
def test_fw_tile128_fused_closure_matches_tile64(gpu_ctx, n):
    """128-wide blocks with at least 4 tiles per CU: each round's remaining-
    tile launch closes the next 128-block (2 x 2 recursion on 64-blocks in
    one workgroup) and the panels run as 64 x 64 quarters.  n = 4200 ends
    with a 104-wide block (both recursion halves), 4500 with a 20-wide one.
    Equal to the 64-block closure (itself checked against the oracle above),
    strided, and sampled rows equal Dijkstra rows of the matrix graph."""
    m = synth.dense_matrix(n, 0.004, 1, 1000, n)
    out = {}
    for tile in (64, 128):
        gpu_ctx.set_option("fw_tile", tile)
        try:
            ld = n + 4 * (tile // 64)
            d = padded(gpu_ctx, m, ld)
            gpu_ctx.fw(d.ptr, n, ld)
            got = d.get().reshape(n, ld)
            d.free()
        finally:
            gpu_ctx.set_option("fw_tile", 0)
        assert (got[:, n:] == -7).all()
        out[tile] = got[:, :n]
    np.testing.assert_array_equal(out[128], out[64])
    rows = np.array([0, 127, 128, n // 2, n - 1])
    mask = (m < O.INF) & ~np.eye(n, dtype=bool)
    src, dst = np.nonzero(mask)
    ro = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(src, minlength=n), out=ro[1:])
    want = O.batched_sssp(rows, ro, dst.astype(np.int32), m[src, dst].astype(np.int32))
    np.testing.assert_array_equal(out[128][rows], want)


def test_fw_golden_and_strided(gpu_ctx):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dense70.npz"))
    for ld in (70, 71, 128):
        d = padded(gpu_ctx, g["mat"], ld)
        gpu_ctx.fw(d.ptr, 70, ld)
        got = d.get().reshape(70, ld)
        np.testing.assert_array_equal(got[:, :70], g["closure"])
        assert (got[:, 70:] == -7).all()


@pytest.mark.parametrize("fw_config", FW_CONFIGS[:2], indirect=True, ids=lambda c: f"T{c[0]}")
def test_fw_positive_diagonal_and_zero_weights(gpu_ctx, fw_config):
    n = 200
    m = synth.dense_matrix(n, 0.03, 0, 50, 77)
    rng = np.random.default_rng(3)
    m[np.arange(n), np.arange(n)] = rng.integers(0, 500, n)  # classical FW keeps min(d_ii, cycle)
    d = gpu_ctx.to_device(m)
    gpu_ctx.fw(d.ptr, n, n)
    np.testing.assert_array_equal(d.get(), O.floyd(m))


# ---------------- SSSP ----------------

@pytest.mark.parametrize("name", ["grid8.npz", "rgg100.npz"])
def test_sssp_golden_all_sources(gpu_ctx, name):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name))
    ro, col, w = g["row_offset"], g["col"], g["wt"]
    V = len(ro) - 1
    dg = gpu_ctx.upload_graph(ro, col, w)
    src = gpu_ctx.to_device(np.arange(V, dtype=np.int32))
    out = gpu_ctx.alloc((V, V))
    dg.sssp(src.ptr, V, out.ptr, V)
    np.testing.assert_array_equal(out.get(), g["apsp"])
    dg.close()


@pytest.fixture
def sssp_config(gpu_ctx, request):
    # (threads, delta, group[, pack]); pack: 1 packed 64-bit words (default),
    # 2 in-row 32-bit words
    threads, delta, group = request.param[:3]
    pack = request.param[3] if len(request.param) > 3 else 1
    gpu_ctx.set_option("sssp_threads", threads)
    gpu_ctx.set_option("sssp_delta", delta)
    gpu_ctx.set_option("sssp_group", group)
    gpu_ctx.set_option("sssp_pack", pack)
    yield request.param
    gpu_ctx.set_option("sssp_threads", 0)
    gpu_ctx.set_option("sssp_delta", 0)
    gpu_ctx.set_option("sssp_group", 0)
    gpu_ctx.set_option("sssp_pack", 0)


def _cfg_id(c):
    return f"t{c[0]}-d{c[1]}-g{c[2]}" + (f"-p{c[3]}" if len(c) > 3 else "")


SSSP_CONFIGS = [(1024, 0, 1), (256, 1, 1), (512, 7, 1), (1024, 100000, 1),
                (1024, 0, 1, 2), (256, 1, 1, 2), (512, 7, 1, 2), (1024, 100000, 1, 2),
                (1024, 0, 8), (256, 1, 8), (512, 7, 4), (1024, 100000, 4), (256, 0, 4),
                (512, 0, 8)]


@pytest.mark.parametrize("sssp_config", SSSP_CONFIGS, indirect=True, ids=_cfg_id)
def test_sssp_grid_sources_ld_and_duplicates(gpu_ctx, sssp_config):
    u, v, w = synth.grid_graph(64, 64, 0.05, 0.02, 1, 60, 1)
    V = 4096
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    sources = np.array([0, 4095, 2080, 17, 17, 3000] + list(range(100, 4000, 97)), np.int32)
    S, ld = len(sources), V + 5
    dsrc = gpu_ctx.to_device(sources)
    out = gpu_ctx.to_device(np.full((S, ld), -7, np.int32))
    dg.sssp(dsrc.ptr, S, out.ptr, ld)
    got = out.get()
    np.testing.assert_array_equal(got[:, :V], O.batched_sssp(sources, ro, col, wt))
    assert (got[:, V:] == -7).all()
    dg.close()


@pytest.mark.parametrize("pack,arcs", [(1, 4), (1, 2), (1, 1), (2, 4)])
@pytest.mark.parametrize("threads,delta,binq", [(256, 0, 1), (256, 1000000, 3), (1024, 0, 2),
                                                (512, 40, 3), (1024, 0, 0)])
def test_sssp_binned_queues(gpu_ctx, threads, delta, binq, pack, arcs):
    """Queues reordered by vertex block (option sssp_bin, each threshold, and
    off) on a 200x200 grid, where near queues exceed the thresholds; both
    state layouts (option sssp_pack) and arcs per relaxation step
    (option sssp_arcs)."""
    u, v, w = synth.grid_graph(200, 200, 0.05, 0.02, 1, 60, 12)
    V = 40000
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    sources = np.array([0, 39999, 20100, 777, 777, 15000], np.int32)
    dsrc = gpu_ctx.to_device(sources)
    out = gpu_ctx.alloc((len(sources), V))
    gpu_ctx.set_option("sssp_threads", threads)
    gpu_ctx.set_option("sssp_delta", delta)
    gpu_ctx.set_option("sssp_bin", binq)
    gpu_ctx.set_option("sssp_pack", pack)
    gpu_ctx.set_option("sssp_arcs", arcs)
    try:
        dg.sssp(dsrc.ptr, len(sources), out.ptr, V)
    finally:
        gpu_ctx.set_option("sssp_bin", 3)
        gpu_ctx.set_option("sssp_pack", 0)
        gpu_ctx.set_option("sssp_arcs", 0)
        gpu_ctx.set_option("sssp_threads", 0)
        gpu_ctx.set_option("sssp_delta", 0)
    np.testing.assert_array_equal(out.get(), O.batched_sssp(sources, ro, col, wt))
    dg.close()


@pytest.mark.parametrize("sssp_config", [(1024, 0, 8), (256, 3, 4)], indirect=True,
                         ids=_cfg_id)
@pytest.mark.parametrize("name", ["grid8.npz", "rgg100.npz"])
def test_sssp_grouped_golden(gpu_ctx, sssp_config, name):
    """Grouped sources (several per workgroup pass), against the golden rows;
    the source list has a ragged last group and repeats inside a group."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name))
    ro, col, w = g["row_offset"], g["col"], g["wt"]
    V = len(ro) - 1
    src = np.concatenate([np.arange(V), [0, 0, V - 1]]).astype(np.int32)
    dg = gpu_ctx.upload_graph(ro, col, w)
    dsrc = gpu_ctx.to_device(src)
    out = gpu_ctx.alloc((len(src), V))
    dg.sssp(dsrc.ptr, len(src), out.ptr, V)
    np.testing.assert_array_equal(out.get(), g["apsp"][src])
    dg.close()


@pytest.mark.parametrize("sssp_config", [(1024, 0, 1), (256, 0, 1, 2), (512, 0, 8)],
                         indirect=True, ids=_cfg_id)
@pytest.mark.parametrize("arcs", [4, 1])
def test_sssp_directed_zero_weights(gpu_ctx, sssp_config, arcs):
    au, av, aw = synth.rgg_graph(500, 5, 0, 9, 5)
    ro, col, w = synth.to_csr(500, au, av, aw, True)
    dg = gpu_ctx.upload_graph(ro, col, w)
    sources = np.arange(0, 500, 3, dtype=np.int32)
    dsrc = gpu_ctx.to_device(sources)
    out = gpu_ctx.alloc((len(sources), 500))
    gpu_ctx.set_option("sssp_arcs", arcs)
    try:
        dg.sssp(dsrc.ptr, len(sources), out.ptr, 500)
    finally:
        gpu_ctx.set_option("sssp_arcs", 0)
    np.testing.assert_array_equal(out.get(), O.batched_sssp(sources, ro, col, w))
    dg.close()


def test_sssp_rejects_out_of_range_source(gpu_ctx):
    import fastapsp_amd as fa
    ro = np.array([0, 1, 2], np.int32)
    dg = gpu_ctx.upload_graph(ro, np.array([1, 0], np.int32), np.array([3, 3], np.int32))
    dsrc = gpu_ctx.to_device(np.array([0, 2], np.int32))
    out = gpu_ctx.alloc((2, 2))
    with pytest.raises(fa.FapError):
        dg.sssp(dsrc.ptr, 2, out.ptr, 2)
    dg.close()


# ---------------- glue + one part end to end ----------------

def partitioned_grid(R, K, seed, wl=1, wh=30):
    u, v, w = synth.grid_graph(R, R, 0.1, 0.03, wl, wh, seed)
    V = R * R
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    gid = (np.arange(V) * K // V + 1).astype(np.int32)
    return V, ro, col, wt, gid, O.precondition(K, gid, ro, col)


def test_gather_scatter_match_oracle(gpu_ctx):
    V, ro, col, wt, gid, pc = partitioned_grid(24, 4, 9)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    for pid in (1, 2):
        s, n, b = int(pc["ver_off"][pid]), int(pc["ver_num"][pid]), int(pc["bdy_num"][pid])
        rb = O.batched_sssp(pc["st2ed"][s:s + b], ro, col, wt)
        drb = gpu_ctx.to_device(rb)
        ld_sub = n + 3
        dsub = gpu_ctx.to_device(np.full((n, ld_sub), -7, np.int32))
        dg.gather(s, n, b, drb.ptr, V, dsub.ptr, ld_sub)
        sub = dsub.get().reshape(n, ld_sub)
        want = O.submat_build(rb, gid, s, n, b, pc["st2ed"], pc["ed2st"], ro, col, wt)
        np.testing.assert_array_equal(sub[:, :n], want)
        assert (sub[:, n:] == -7).all()
        blk0 = rand_mat(n, V, 55)
        dblk = gpu_ctx.to_device(blk0)
        dg.scatter(s, n, n, dsub.ptr, ld_sub, dblk.ptr, V)
        np.testing.assert_array_equal(dblk.get(), O.submat_decode(want, blk0, s, pc["st2ed"]))
    dg.close()


@pytest.mark.parametrize("R,K", [(20, 4), (33, 7), (16, 1)])
def test_subgraph_solve_matches_oracle(gpu_ctx, R, K):
    V, ro, col, wt, gid, pc = partitioned_grid(R, K, R + K)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    for pid in range(1, K + 1):
        s, n, b = int(pc["ver_off"][pid]), int(pc["ver_num"][pid]), int(pc["bdy_num"][pid])
        out = gpu_ctx.alloc((n, V))
        dg.subgraph_solve(s, n, b, out.ptr, V)
        got = out.get()
        np.testing.assert_array_equal(got, O.run_part(pid, ro, col, wt, gid, pc))
        out.free()
    dg.close()


@pytest.fixture(params=[1, 0], ids=["index", "sssp"])
def stage1(gpu_ctx, request):
    """Stage 1 of batched part solves through the boundary index (default) or
    the SSSP over the whole graph."""
    gpu_ctx.set_option("sssp_index", request.param)
    yield request.param
    gpu_ctx.set_option("sssp_index", 1)


@pytest.mark.parametrize("sssp_config", [(1024, 0, 0), (1024, 0, 0, 2), (512, 0, 8)],
                         indirect=True, ids=_cfg_id)
@pytest.mark.parametrize("R,K,pick", [(20, 4, None), (33, 7, (7, 2, 5)), (16, 1, None),
                                       (40, 12, (3, 3, 12, 1))])
def test_subgraph_solve_batch_matches_per_part(gpu_ctx, stage1, sssp_config, R, K, pick):
    """All parts (or a subset, duplicates included) in one batched call: every
    row block equals the oracle's per-part pipeline, and a guard column past
    each block's V columns is left alone."""
    V, ro, col, wt, gid, pc = partitioned_grid(R, K, 3 * R + K)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    ids = list(pick) if pick else list(range(1, K + 1))
    ld = V + 5
    outs = [gpu_ctx.to_device(np.full((int(pc["ver_num"][p]), ld), -3, np.int32)) for p in ids]
    dg.subgraph_solve_batch([int(pc["ver_off"][p]) for p in ids],
                            [int(pc["ver_num"][p]) for p in ids],
                            [int(pc["bdy_num"][p]) for p in ids],
                            [o.ptr for o in outs], ld)
    for p, o in zip(ids, outs):
        got = o.get().reshape(-1, ld)
        np.testing.assert_array_equal(got[:, :V], O.run_part(p, ro, col, wt, gid, pc))
        assert (got[:, V:] == -3).all()
        o.free()
    dg.close()


def _rgg_partitioned(n, K, seed, wl, wh):
    u, v, w = synth.rgg_graph(n, 6, wl, wh, seed)
    ro, col, wt = synth.to_csr(n, u, v, w, True)
    # parts by x coordinate strips of the vertex ids (RGG ids are random)
    gid = (np.arange(n) * K // n + 1).astype(np.int32)
    return ro, col, wt, gid, O.precondition(K, gid, ro, col)


@pytest.mark.parametrize("n,K,wl,wh", [(900, 6, 1, 255), (1500, 9, 0, 3), (400, 1, 1, 9)])
def test_boundary_index_rows_equal_sssp_rows(gpu_ctx, n, K, wl, wh):
    """Directed random geometric graphs (weights 0..3 give equal-distance
    plateaus; K = 1 has no boundary at all): the batched solve with stage 1
    through the boundary index equals the one through the whole-graph SSSP
    and the oracle, part by part."""
    ro, col, wt, gid, pc = _rgg_partitioned(n, K, 60 + n, wl, wh)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    assert dg.build_index() == (K > 1)           # no boundary vertices: no index needed
    nb, nbytes = dg.index_info()
    assert nb == int(pc["bdy_num"][1:].sum()) and (nbytes > 0) == (K > 1)
    ids = list(range(1, K + 1))
    got = {}
    # stage 1 through the index with 64x128 and 64x64 block tiles and a
    # narrow band, and through SSSP
    for cfg in ((1, 0, 0), (1, 64, 0), (1, 128, 50), (0, 0, 0)):
        gpu_ctx.set_option("sssp_index", cfg[0])
        gpu_ctx.set_option("bidx_tile", cfg[1])
        gpu_ctx.set_option("bidx_delta", cfg[2])
        try:
            outs = [gpu_ctx.alloc((int(pc["ver_num"][p]), n)) for p in ids]
            dg.subgraph_solve_batch([int(pc["ver_off"][p]) for p in ids],
                                    [int(pc["ver_num"][p]) for p in ids],
                                    [int(pc["bdy_num"][p]) for p in ids], [o.ptr for o in outs], n)
            got[cfg] = [o.get() for o in outs]
            for o in outs:
                o.free()
        finally:
            gpu_ctx.set_option("sssp_index", 1)
            gpu_ctx.set_option("bidx_tile", 0)
            gpu_ctx.set_option("bidx_delta", 0)
    for i, p in enumerate(ids):
        want = O.run_part(p, ro, col, wt, gid, pc)
        for cfg, rows in got.items():
            np.testing.assert_array_equal(rows[i], want, err_msg=f"part {p}, {cfg}")
    dg.close()


@pytest.mark.parametrize("index", [1, 0])
def test_subgraph_boundary_rows_match_dijkstra(gpu_ctx, index):
    """Stage 1 alone (fap_subgraph_boundary_rows, the reference's
    handle_boundry_path for several parts): rows 0..b-1 of each part's row
    block equal Dijkstra from the part's boundary vertices, through the
    boundary index and through SSSP; the other rows are left alone."""
    ro, col, wt, gid, pc = _rgg_partitioned(1100, 7, 91, 0, 40)
    n = len(ro) - 1
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    ids = [2, 3, 6]
    gpu_ctx.set_option("sssp_index", index)
    try:
        outs = [gpu_ctx.alloc((int(pc["ver_num"][p]), n)) for p in ids]
        for o in outs:
            gpu_ctx.fill(o, -5)
        dg.boundary_rows([int(pc["ver_off"][p]) for p in ids], [int(pc["ver_num"][p]) for p in ids],
                         [int(pc["bdy_num"][p]) for p in ids], [o.ptr for o in outs], n)
        for p, o in zip(ids, outs):
            b, s0 = int(pc["bdy_num"][p]), int(pc["ver_off"][p])
            got = o.get()
            np.testing.assert_array_equal(got[:b], O.batched_sssp(pc["st2ed"][s0:s0 + b], ro, col, wt))
            assert (got[b:] == -5).all()
            o.free()
    finally:
        gpu_ctx.set_option("sssp_index", 1)
    dg.close()


@pytest.mark.parametrize("R,K,ids", [(64, 96, None), (150, 2100, (1, 700, 1400, 2100))])
def test_boundary_index_scheduling_beyond_lds_tables(gpu_ctx, R, K, ids):
    """The scheduling step keeps block marks in LDS up to 8,192 blocks (parts
    in the call x K) and part tables up to 2,048 parts; beyond, it works from
    global memory.  96 parts in one call (9,216 blocks) and a 2,100-part
    graph (both beyond): every part's rows equal those of the SSSP stage and
    of the oracle."""
    V, ro, col, wt, gid, pc = partitioned_grid(R, K, R + K)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    assert dg.build_index()
    ids = list(ids) if ids else list(range(1, K + 1))
    sums = {}
    for flag in (1, 0):
        gpu_ctx.set_option("sssp_index", flag)
        try:
            outs = [gpu_ctx.alloc((int(pc["ver_num"][p]), V)) for p in ids]
            dg.subgraph_solve_batch([int(pc["ver_off"][p]) for p in ids],
                                    [int(pc["ver_num"][p]) for p in ids],
                                    [int(pc["bdy_num"][p]) for p in ids], [o.ptr for o in outs], V)
            sums[flag] = [gpu_ctx.checksum(o.ptr, int(pc["ver_num"][p]) * V)
                          for o, p in zip(outs, ids)]
            if flag == 1:
                for p, o in list(zip(ids, outs))[:: max(1, len(ids) // 3)]:
                    np.testing.assert_array_equal(o.get(), O.run_part(p, ro, col, wt, gid, pc),
                                                  err_msg=f"part {p}")
            for o in outs:
                o.free()
        finally:
            gpu_ctx.set_option("sssp_index", 1)
    assert sums[1] == sums[0]
    dg.close()


def test_boundary_index_survives_staging_growth_between_calls():
    """Regression: a batched solve whose descriptor staging outgrows the
    previous call's (40 parts, ~12,000 sources, after a one-part call) must
    leave the boundary iterations' own pinned buffer alone.  A fresh context
    runs one part, then all 40, then one again; every part equals the
    whole-graph SSSP path (checksums), the two end parts the oracle."""
    V, ro, col, wt, gid, pc = partitioned_grid(160, 40, 21)
    K = 40
    ctx = fa.Context(0)
    try:
        dg = ctx.upload_graph(ro, col, wt)
        dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
        assert dg.build_index()
        print(f"[bidx staging] V={V} K={K} sources={int(pc['bdy_num'][1:].sum())}", flush=True)
        outs = {p: ctx.alloc((int(pc["ver_num"][p]), V)) for p in range(1, K + 1)}

        def solve(ids):
            dg.subgraph_solve_batch([int(pc["ver_off"][p]) for p in ids],
                                    [int(pc["ver_num"][p]) for p in ids],
                                    [int(pc["bdy_num"][p]) for p in ids],
                                    [outs[p].ptr for p in ids], V)
            return {p: ctx.checksum(outs[p].ptr, int(pc["ver_num"][p]) * V) for p in ids}

        first = solve([3])
        every = solve(list(range(1, K + 1)))
        last = solve([5])
        for p in (1, 40):
            np.testing.assert_array_equal(outs[p].get(), O.run_part(p, ro, col, wt, gid, pc),
                                          err_msg=f"part {p}")
        ctx.set_option("sssp_index", 0)
        plain = solve(list(range(1, K + 1)))
        assert every == plain
        assert first[3] == plain[3] and last[5] == plain[5]
        for o in outs.values():
            o.free()
        dg.close()
    finally:
        ctx.close()


def test_boundary_index_declines_duplicate_arcs(gpu_ctx):
    """A graph with a duplicate arc gets no index (its part matrix keeps the
    last duplicate, as the reference's does); batched solves then run the
    whole-graph SSSP and still match the oracle."""
    V, ro, col, wt, gid, pc = partitioned_grid(14, 3, 5)
    ro2 = ro.copy()
    col2 = np.insert(col, ro[3], col[ro[3]])      # vertex 3's first arc twice
    wt2 = np.insert(wt, ro[3], wt[ro[3]] + 7)
    ro2[4:] += 1
    dg = gpu_ctx.upload_graph(ro2, col2, wt2)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    assert not dg.build_index()
    assert dg.index_info() == (0, 0)
    outs = [gpu_ctx.alloc((int(pc["ver_num"][p]), V)) for p in (1, 2, 3)]
    dg.subgraph_solve_batch([int(pc["ver_off"][p]) for p in (1, 2, 3)],
                            [int(pc["ver_num"][p]) for p in (1, 2, 3)],
                            [int(pc["bdy_num"][p]) for p in (1, 2, 3)], [o.ptr for o in outs], V)
    for p, o in zip((1, 2, 3), outs):
        np.testing.assert_array_equal(o.get(), O.run_part(p, ro2, col2, wt2, gid, pc))
        o.free()
    dg.close()


def test_subgraph_solve_batch_rejects_bad_parts(gpu_ctx):
    V, ro, col, wt, gid, pc = partitioned_grid(12, 3, 1)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    out = gpu_ctx.alloc((V, V))
    s, n, b = int(pc["ver_off"][2]), int(pc["ver_num"][2]), int(pc["bdy_num"][2])
    with pytest.raises(fa.FapError):
        dg.subgraph_solve_batch([s], [n], [n + 1], [out.ptr], V)   # n_bdy > n_sub
    with pytest.raises(fa.FapError):
        dg.subgraph_solve_batch([s], [n], [b], [out.ptr], V - 1)   # ld < V
    with pytest.raises(fa.FapError):
        dg.subgraph_solve_batch([s + 1], [n - 1], [b], [out.ptr], V)   # not a part start
    with pytest.raises(fa.FapError):
        dg.subgraph_solve_batch([s], [n + 1], [b], [out.ptr], V)   # runs into the next part
    dg.subgraph_solve_batch([], [], [], [], V)                     # empty batch is fine
    out.free()
    dg.close()


@pytest.mark.parametrize("group_parts", [1, 3, 5, 100])
def test_subgraph_solve_groups_matches_per_part(gpu_ctx, group_parts):
    """Parts in groups on two streams: every row block equals the oracle's."""
    R, K = 36, 11
    V, ro, col, wt, gid, pc = partitioned_grid(R, K, 77)
    dg = gpu_ctx.upload_graph(ro, col, wt)
    dg.set_partition(gid, pc["st2ed"], pc["ed2st"])
    ids = list(range(1, K + 1))
    outs = [gpu_ctx.alloc((int(pc["ver_num"][p]), V)) for p in ids]
    dg.subgraph_solve_groups([int(pc["ver_off"][p]) for p in ids],
                             [int(pc["ver_num"][p]) for p in ids],
                             [int(pc["bdy_num"][p]) for p in ids],
                             [o.ptr for o in outs], V, group_parts)
    for p, o in zip(ids, outs):
        np.testing.assert_array_equal(o.get(), O.run_part(p, ro, col, wt, gid, pc))
        o.free()
    dg.close()
