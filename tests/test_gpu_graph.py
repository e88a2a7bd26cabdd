"""The fap::fapGraph drop-in end to end on the GPU, used the way the
reference's example uses it (app/singleNodeExample.cpp:40-68), checked
bit-exact against the oracle (the distances from every vertex of a part,
which do not depend on the partition)."""

import os

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth, ops

pytestmark = pytest.mark.gpu


def grid_csr(R, drop, diag, wl, wh, seed):
    u, v, w = synth.grid_graph(R, R, drop, diag, wl, wh, seed)
    V = R * R
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    return V, u, v, w, ro, col, wt


def test_plain_small_full_matrix():
    """64x64 grid, 5% of edges removed, unit weights, K=8: solve() gives the
    full 4096 x 4096 matrix (rows in st2ed order, columns vertex ids)."""
    V, u, v, w, ro, col, wt = grid_csr(64, 0.05, 0.0, 1, 1, 1)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), 8) as g:
        g.preCondition()
        assert not g.isSplit()
        g.solve()
        D = g.getAllDistanceInt32()
        st2ed = g.getMapping()
        want = O.batched_sssp(st2ed, ro, col, wt)
        np.testing.assert_array_equal(D, want)
        F = g.getAllDistance()
        np.testing.assert_array_equal(O.from_ref(F), want)
        assert g.getCurrentSubGraphId() == 8
        # the example's verification on the last part
        assert g.check_result(8)


def test_file_constructor_matches_reference_reader(tmp_path, monkeypatch):
    V, u, v, w, ro, col, wt = grid_csr(20, 0.1, 0.05, 1, 50, 5)
    d = tmp_path / "graph" / "weight-undirected"
    synth.write_mtx(str(d / "g20.mtx"), V, u, v, w)
    monkeypatch.chdir(tmp_path)
    with fa.Graph.from_file("g20", False, True, 4) as g:
        r_ro, r_col, r_w = g.csr()
        e_ro, e_col, e_w = synth.csr_reference_order(V, u, v, w, False)
        np.testing.assert_array_equal(r_ro, e_ro)
        np.testing.assert_array_equal(r_col, e_col)
        np.testing.assert_array_equal(r_w, e_w)
        g.preCondition()
        g.solve()
        np.testing.assert_array_equal(g.getAllDistanceInt32(),
                                      O.batched_sssp(g.getMapping(), ro, col, wt))


def test_split_graph_part_by_part():
    """V > 20000 -> isSplit(); parts solved one at a time, as the example does."""
    V, u, v, w, ro, col, wt = grid_csr(150, 0.1, 0.02, 1, 100, 3)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), 24) as g:
        g.preCondition()
        assert g.isSplit()
        vn, vo, bn, _ = g.part_info()
        for pid in (1, 2, 3, 24):
            g.solveSubGraph(pid)
            assert g.getCurrentSubGraphId() == pid
            idx = g.getSubGraphIndex(pid)
            blk = g.getSubGraphDistanceInt32(pid)
            assert blk.shape == (vn[pid], V)
            np.testing.assert_array_equal(blk, O.batched_sssp(idx, ro, col, wt))
            # point queries and one-row getter on the current part
            x = int(idx[len(idx) // 2])
            row = g.getDistanceFromOnePoint(x)
            np.testing.assert_array_equal(O.from_ref(row), blk[len(idx) // 2])
            assert g.getDistanceP2P(x, 0) == blk[len(idx) // 2, 0]
            s, ws = g.checksum(pid)
            assert s == int(blk.astype(np.uint32).astype(np.uint64).sum()) % (1 << 64)
        assert g.check_result(24)


def test_directed_rgg_graph():
    au, av, aw = synth.rgg_graph(700, 6, 1, 255, 9)
    ro, col, w = synth.to_csr(700, au, av, aw, True)
    # both arcs of every pair, independent weights: the distances are not symmetric
    with fa.Graph.from_csr(ro, col, w.astype(np.float32), 5) as g:
        g.preCondition()
        g.solve()
        np.testing.assert_array_equal(g.getAllDistanceInt32(),
                                      O.batched_sssp(g.getMapping(), ro, col, w))


def test_errors_are_reported_not_fatal():
    V, u, v, w, ro, col, wt = grid_csr(10, 0.0, 0.0, 1, 1, 1)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), 3) as g:
        with pytest.raises(fa.FapError):
            g.solve()                 # before preCondition
        g.preCondition()
        with pytest.raises(fa.FapError):
            g.solveSubGraph(4)        # part id outside 1..K
        with pytest.raises(fa.FapError):
            g.getSubGraphDistance(2)  # not solved
        g.solveSubGraph(2)
        assert g.getSubGraphDistanceInt32(2).shape[1] == V


# ---------------- the reference's float operators ----------------

def test_ops_float_mirror():
    m = synth.dense_matrix(150, 0.05, 1, 1000, 8)
    f = O.to_ref(m)
    ops.floyd_path(150, f)
    np.testing.assert_array_equal(O.from_ref(f), O.floyd(m))
    B = np.ascontiguousarray(O.to_ref(synth.dense_matrix(40, 0.5, 1, 99, 2)))
    res = np.empty((60, 40), np.float32)
    A = np.ascontiguousarray(O.to_ref(synth.dense_matrix(60, 0.5, 1, 99, 1))[:, :40])
    ops.min_plus_path_advanced(A, B, res, 60, 40, 40)
    np.testing.assert_array_equal(O.from_ref(res), O.min_plus(O.from_ref(A), O.from_ref(B)))
    V, u, v, w, ro, col, wt = grid_csr(30, 0.1, 0.0, 1, 20, 4)
    st2ed = np.arange(V, dtype=np.int32)[::-1].copy()
    sub = np.zeros((5, V), np.float32)
    ops.handle_boundry_path(sub, V, len(col), 5, ro, col, wt.astype(np.float32), st2ed, 10)
    np.testing.assert_array_equal(O.from_ref(sub), O.batched_sssp(st2ed[10:15], ro, col, wt))


def test_batched_and_plain_solve_agree():
    """solve() runs the batched form by default; the plain part loop and the
    oracle give the same full matrix.  solveSubGraphs keeps several parts'
    rows available at once."""
    V, u, v, w, ro, col, wt = grid_csr(48, 0.08, 0.03, 1, 40, 11)
    want = None
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), 9) as g:
        g.preCondition()
        g.solve()
        D = g.getAllDistanceInt32()
        want = O.batched_sssp(g.getMapping(), ro, col, wt)
        np.testing.assert_array_equal(D, want)
        g.setBatched(False)
        g.solve()
        np.testing.assert_array_equal(g.getAllDistanceInt32(), want)
        vn, vo, _, _ = g.part_info()
        for batched in (True, False):
            g.setBatched(batched)
            g.solveSubGraphs([4, 9, 1, 4])
            assert g.getCurrentSubGraphId() == 4
            for pid in (1, 4, 9):
                np.testing.assert_array_equal(g.getSubGraphDistanceInt32(pid),
                                              want[vo[pid]:vo[pid] + vn[pid]])


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_documented_run_luxembourg_k128(tmp_path, monkeypatch):
    """The reference's documented run, `singleNodeExample -f luxembourg_osm
    -k 128 -direct false -weight false` (README.md:57, app/README.md:3,
    app/singleNodeExample.cpp:40-68), on its own fixture graph, replayed
    byte for byte from tests/golden/luxembourg_osm.npz into the reference's
    directory layout: fapGraph from the file, preCondition, split, parts 1-3
    solved with paths, the example's check_result on the current part (3),
    sampled rows of every part equal to Dijkstra, and part 3's predecessors
    valid under the reference's path semantics."""
    z = np.load(os.path.join(GOLDEN, "luxembourg_osm.npz"))
    V, lines, edges = int(z["V"]), int(z["lines"]), z["edges"]
    d = tmp_path / "graph" / "unweight-undirected"
    d.mkdir(parents=True)
    with open(d / "luxembourg_osm.mtx", "w") as f:
        f.write(f"{V} {V} {lines}\n")
        f.write("".join(f"{a} {b}\n" for a, b in edges))
    monkeypatch.chdir(tmp_path)
    with fa.Graph.from_file("luxembourg_osm", False, False, 128) as g:
        assert g.num_vertexs == V
        ro, col, wt = g.csr()
        g.preCondition()
        assert g.isSplit()
        rng = np.random.default_rng(128)
        for pid in (1, 2, 3):            # each solve replaces the previous part's rows
            g.solveSubGraph(pid, True)
            idx = g.getSubGraphIndex(pid)
            n = len(idx)
            rs = np.unique(np.r_[0, n - 1, rng.integers(0, n, 4)]).astype(np.int32)
            want = O.batched_sssp(idx[rs], ro, col, wt)
            for k, r in enumerate(rs):
                np.testing.assert_array_equal(g.getSubGraphRowsInt32(pid, int(r), 1)[0], want[k],
                                              err_msg=f"part {pid} row {r}")
        assert g.getCurrentSubGraphId() == 3
        assert g.check_result(3)
        D, P = g.getSubGraphDistanceInt32(3), g.getSubGraphPath(3)
        assert O.check_pred(D, P, g.getSubGraphIndex(3), ro, col, wt) == 0
