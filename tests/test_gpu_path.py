"""Path rows (is_path_needed) on the GPU.  The reference's path matrices
(fap/kernel/batch_sssp.h:82-83,110; fap/kernel/floyd.h:44;
fap/kernel/minplus.h:57-60; fap/utils/util_Myalgorithm.h:50-51,71) hold the
vertex before each column on a shortest path; the reference's own backends
break ties differently, so rows are checked for validity with the oracle's
checker (pinned against the reference's Dijkstra and run() paths in
test_oracle.py), distances bit-exact as everywhere."""

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth

pytestmark = pytest.mark.gpu


def grid(R, seed, wl=1, wh=50):
    u, v, w = synth.grid_graph(R, R, 0.1, 0.03, wl, wh, seed)
    return synth.to_csr(R * R, u, v, w, False)


def walk_cost(path, ro, col, w):
    """Sum of the cheapest arc weights along a vertex sequence."""
    tot = 0
    for a, b in zip(path[:-1], path[1:]):
        ws = [w[e] for e in range(ro[a], ro[a + 1]) if col[e] == b]
        assert ws, (a, b)
        tot += min(ws)
    return tot


@pytest.mark.parametrize("batched", [True, False])
def test_solve_with_paths_full_matrix(batched):
    ro, col, w = grid(40, 7)
    with fa.Graph.from_csr(ro, col, w.astype(np.float32), 6) as g:
        g.setBatched(batched)
        g.preCondition()
        g.solve(True)
        D = g.getAllDistanceInt32()
        P = g.getAllPath()
        own = g.getMapping()
        np.testing.assert_array_equal(D, O.batched_sssp(own, ro, col, w))
        assert O.check_pred(D, P, own, ro, col, w) == 0
        # P2P: a shortest path whose arcs add up to the distance
        for src, dst in ((own[0], own[-1]), (own[100], own[700]), (5, 5)):
            seq = g.getPathP2P(int(src), int(dst))
            assert seq[0] == src and seq[-1] == dst
            assert walk_cost(seq, ro, col, w) == D[np.flatnonzero(own == src)[0], dst]


def test_solve_subgraph_paths_split_graph():
    ro, col, w = grid(150, 3, 1, 100)
    with fa.Graph.from_csr(ro, col, w.astype(np.float32), 24) as g:
        g.preCondition()
        assert g.isSplit()
        for pid in (1, 13):
            g.solveSubGraph(pid, True)
            idx = g.getSubGraphIndex(pid)
            D = g.getSubGraphDistanceInt32(pid)
            P = g.getSubGraphPath(pid)
            assert O.check_pred(D, P, idx, ro, col, w) == 0
            x = int(idx[3])
            np.testing.assert_array_equal(g.getPathFromOnePoint(x), P[3])
        g.solveSubGraphs([2, 24], True)
        for pid in (2, 24):
            assert O.check_pred(g.getSubGraphDistanceInt32(pid), g.getSubGraphPath(pid),
                                g.getSubGraphIndex(pid), ro, col, w) == 0
        g.solveSubGraph(5)                       # without paths: path getters refuse
        with pytest.raises(fa.FapError):
            g.getSubGraphPath(5)


def test_paths_with_zero_weight_plateaus(gpu_ctx):
    """Directed RGG with weights 0..3: many zero-weight arcs and equal-distance
    plateaus, resolved without cycles."""
    au, av, aw = synth.rgg_graph(600, 7, 0, 3, 12)
    ro, col, w = synth.to_csr(600, au, av, aw, True)
    dg = gpu_ctx.upload_graph(ro, col, w)
    src = np.arange(0, 600, 5, dtype=np.int32)
    dsrc = gpu_ctx.to_device(src)
    D = gpu_ctx.alloc((len(src), 600))
    dg.sssp(dsrc.ptr, len(src), D.ptr, 600)
    P = gpu_ctx.to_device(np.full((len(src), 603), -9, np.int32))
    dg.pred_rows(D.ptr, 600, dsrc.ptr, len(src), P.ptr, 603)
    d, p = D.get(), P.get().reshape(len(src), 603)
    np.testing.assert_array_equal(d, O.batched_sssp(src, ro, col, w))
    assert O.check_pred(d, p[:, :600], src, ro, col, w) == 0
    assert (p[:, 600:] == -9).all()
    # plateau passes only use cells resolved in earlier passes: the same rows again
    P2 = gpu_ctx.to_device(np.full((len(src), 603), -9, np.int32))
    dg.pred_rows(D.ptr, 600, dsrc.ptr, len(src), P2.ptr, 603)
    np.testing.assert_array_equal(P2.get().reshape(len(src), 603), p)
    dg.close()


def test_pred_rows_rejects_bad_arguments(gpu_ctx):
    ro, col, w = grid(8, 1)
    dg = gpu_ctx.upload_graph(ro, col, w)
    D = gpu_ctx.alloc((2, 64))
    with pytest.raises(fa.FapError):
        dg.pred_rows(D.ptr, 63, D.ptr, 2, D.ptr, 64)       # ld < V
    with pytest.raises(fa.FapError):
        dg.subgraph_pred(0, 2, D.ptr, 64, D.ptr, 64)       # no partition attached
    dg.close()


# ---------------- path outputs of the operator boundary ----------------

def _ties(rng, shape, hi, p_inf):
    a = rng.integers(0, hi, shape).astype(np.int32)
    a[rng.random(shape) < p_inf] = O.INF
    return a


@pytest.mark.parametrize("M,N,K", [(130, 257, 70), (64, 64, 16), (1, 5, 1), (77, 300, 129),
                                   (200, 100, 0), (65, 1, 33)])
def test_minplus_path_equals_reference_rule(gpu_ctx, M, N, K):
    """fap_minplus_path: distances and the label of the first minimising k,
    bit-exact against the oracle's min_plus_path_CPU_advanced restatement
    (pinned on the reference's own code); small weights force ties; strided
    operands."""
    rng = np.random.default_rng(M * 7 + N + K)
    A, B = _ties(rng, (M, K), 5, 0.2), _ties(rng, (K, N), 5, 0.2)
    if M > 1:
        A[1, :] = O.INF
    Bp = rng.integers(0, 1 << 20, (K, N)).astype(np.int32)
    lda, ldb, ldc = K + 3, N + 5, N + 2
    pad = lambda a, ld: np.pad(a, ((0, 0), (0, ld - a.shape[1])), constant_values=-7)
    dA, dB = gpu_ctx.to_device(pad(A, lda)), gpu_ctx.to_device(pad(B, ldb))
    dBp = gpu_ctx.to_device(pad(Bp, ldb))
    dC = gpu_ctx.to_device(np.full((M, ldc), -9, np.int32))
    dCp = gpu_ctx.to_device(np.full((M, ldc), -9, np.int32))
    gpu_ctx.minplus_path(dA.ptr, lda, dB.ptr, ldb, dBp.ptr, ldb, dC.ptr, ldc, dCp.ptr, ldc, M, N, K)
    gpu_ctx.sync()
    want, wantp = O.min_plus(A, B, Bp)
    got, gotp = dC.get(), dCp.get()
    np.testing.assert_array_equal(got[:, :N], want)
    np.testing.assert_array_equal(gotp[:, :N], wantp)
    assert (got[:, N:] == -9).all() and (gotp[:, N:] == -9).all()


def _fw_input(n, seed, p_arc=None):
    rng = np.random.default_rng(seed)
    m = rng.integers(0, 6, (n, n)).astype(np.int32)     # zero weights: plateaus
    m[rng.random((n, n)) >= (p_arc or min(1.0, 5.0 / max(n, 1)))] = O.INF
    if n > 3:
        m[:, n - 1] = O.INF                              # an unreachable column
    np.fill_diagonal(m, 0)
    return m


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 257])
def test_fw_path_hops_valid_and_labels_carried(gpu_ctx, n):
    """fap_fw_blocked_path: distances equal the oracle's FW; with identity
    labels (the caller's path[i][j] = i for an arc, -1 for none, j on the
    diagonal, as fap/utils/util_Myalgorithm.h:50-51,71 builds them) the hops
    pass the oracle's checker (tight, chains end at the row); with arbitrary
    labels every cell carries the label of the same hop."""
    m = _fw_input(n, 100 + n)
    ident = np.where(m < O.INF, np.arange(n, dtype=np.int32)[:, None], -1).astype(np.int32)
    np.fill_diagonal(ident, np.arange(n))
    ld, ldp = n + 5, n + 3
    pad = lambda a, w: np.pad(a, ((0, 0), (0, w - n)), constant_values=-7)

    def run(labels):
        d = gpu_ctx.to_device(pad(m, ld))
        p = gpu_ctx.to_device(pad(labels, ldp))
        gpu_ctx.fw_path(d.ptr, n, ld, p.ptr, ldp)
        dd, pp = d.get(), p.get()
        assert (dd[:, n:] == -7).all() and (pp[:, n:] == -7).all()
        return dd[:, :n], pp[:, :n]

    d, hop = run(ident)
    np.testing.assert_array_equal(d, O.floyd(m))
    assert O.fw_hop_check(m, d, hop) == 0
    lab = np.random.default_rng(n).integers(0, 1 << 30, (n, n)).astype(np.int32)
    d2, p2 = run(lab)
    np.testing.assert_array_equal(d2, d)
    improved = d < m
    want = np.where(improved, lab[np.where(improved, hop, 0), np.arange(n)[None, :]], lab)
    np.testing.assert_array_equal(p2, want)


def test_fw_path_rejects_bad_arguments(gpu_ctx):
    d = gpu_ctx.alloc((4, 4))
    with pytest.raises(fa.FapError):
        gpu_ctx.fw_path(d.ptr, 4, 3, d.ptr, 4)
    with pytest.raises(fa.FapError):
        gpu_ctx.minplus_path(d.ptr, 4, d.ptr, 4, 0, 4, d.ptr, 4, d.ptr, 4, 4, 4, 4)  # no B_path


def test_ops_float_boundary_writes_paths():
    """The reference's operator boundary (fap_ops.h) with path buffers, as
    fapGraph::run passes them (fap/fap.cpp:101-103,106,123,144): SSSP
    parents, FW hop labels and min-plus labels, all valid (and the min-plus
    labels bit-exact)."""
    from fastapsp_amd import ops
    # stage 1: parents of 6 boundary sources of a weighted grid
    ro, col, w = grid(30, 4, 1, 20)
    V = len(ro) - 1
    st2ed = np.arange(V, dtype=np.int32)[::-1].copy()
    sub = np.zeros((6, V), np.float32)
    subp = np.full((6, V), -5, np.int32)
    ops.handle_boundry_path(sub, V, len(col), 6, ro, col, w.astype(np.float32), st2ed, 10, subp)
    d = O.from_ref(sub)
    np.testing.assert_array_equal(d, O.batched_sssp(st2ed[10:16], ro, col, w))
    assert O.check_pred(d, subp, st2ed[10:16], ro, col, w) == 0
    # stage 2: FW with identity labels
    n = 150
    m = _fw_input(n, 9)
    ident = np.where(m < O.INF, np.arange(n, dtype=np.int32)[:, None], -1).astype(np.int32)
    np.fill_diagonal(ident, np.arange(n))
    f = O.to_ref(m)
    ops.floyd_path(n, f, ident)
    df = O.from_ref(f)
    np.testing.assert_array_equal(df, O.floyd(m))
    assert O.fw_hop_check(m, df, ident) == 0
    # stage 3: min-plus labels, exact
    rng = np.random.default_rng(2)
    A, B = _ties(rng, (60, 40), 4, 0.3), _ties(rng, (40, 90), 4, 0.3)
    Bp = rng.integers(0, V, (40, 90)).astype(np.int32)
    res = np.empty((60, 90), np.float32)
    resp = np.full((60, 90), -5, np.int32)
    ops.min_plus_path_advanced(np.ascontiguousarray(O.to_ref(A)), np.ascontiguousarray(O.to_ref(B)),
                               res, 60, 90, 40, Bp, resp)
    want, wantp = O.min_plus(A, B, Bp)
    np.testing.assert_array_equal(O.from_ref(res), want)
    np.testing.assert_array_equal(resp, wantp)
