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
