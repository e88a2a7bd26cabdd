"""Full-size workloads on the GPU, checked through size-independent
properties and sampled exact rows (the oracle would take minutes on the full
outputs):

* dense n = 8192 (arc p = 0.01, w 1..1000): the closure's rows equal the
  Dijkstra rows of the matrix graph; the min-plus product's sampled rows
  equal the oracle's product rows.
* 512 x 512 road-like grid (-10% edges, +2% diagonals, w 1..100), K = 256:
  part blocks' sampled rows equal Dijkstra rows; symmetry across parts
  (undirected); device checksums equal host sums.
"""

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def matrix_csr(m):
    n = m.shape[0]
    mask = (m < O.INF) & ~np.eye(n, dtype=bool)
    src, dst = np.nonzero(mask)
    ro = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(src, minlength=n), out=ro[1:])
    return ro, dst.astype(np.int32), m[src, dst].astype(np.int32)


def test_dense_8192_closure_and_product(gpu_ctx):
    n = 8192
    m = synth.dense_matrix(n, 0.01, 1, 1000, 2)
    d = gpu_ctx.to_device(m)
    gpu_ctx.fw(d.ptr, n, n)
    rows = np.array([0, 1, 777, 4096, 8191])
    got = d.get()[rows]
    ro, col, w = matrix_csr(m)
    want = O.batched_sssp(rows, ro, col, w)
    np.testing.assert_array_equal(got, want)
    d.free()
    A = synth.dense_matrix(n, 0.01, 1, 1000, 21)
    B = synth.dense_matrix(n, 0.01, 1, 1000, 22)
    dA, dB, dC = gpu_ctx.to_device(A), gpu_ctx.to_device(B), gpu_ctx.alloc((n, n))
    gpu_ctx.minplus(dA.ptr, n, dB.ptr, n, dC.ptr, n, n, n, n)
    C_ = dC.get()
    sel = np.array([3, 5000, 8190])
    np.testing.assert_array_equal(C_[sel], O.min_plus(A[sel], B))
    # both operands have a zero diagonal: C[i][i] = 0, everything else >= 1
    assert (np.diag(C_) == 0).all() and (C_ <= O.INF).all()
    assert (C_ + np.eye(n, dtype=np.int32) >= 1).all()
    for x in (dA, dB, dC):
        x.free()


def test_grid_512_k256_sampled_parts():
    R, K = 512, 256
    V = R * R
    u, v, w = synth.grid_graph(R, R, 0.10, 0.02, 1, 100, 3)
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    assert len(col) == 952362
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        g.preCondition()
        assert g.isSplit()
        vn, vo, bn, _ = g.part_info()
        blocks = {}
        for pid in (1, 129, 256):
            g.solveSubGraph(pid)
            idx = g.getSubGraphIndex(pid)
            blk = g.getSubGraphDistanceInt32(pid)
            # boundary rows, inner rows, the last row
            pick = np.unique([0, int(bn[pid]) - 1, int(bn[pid]), len(idx) // 2, len(idx) - 1])
            np.testing.assert_array_equal(blk[pick], O.batched_sssp(idx[pick], ro, col, wt))
            s, _ = g.checksum(pid)
            assert s == int(blk.astype(np.uint64).sum()) % (1 << 64)
            assert (blk[np.arange(len(idx)), idx] == 0).all()
            blocks[pid] = (idx, blk)
        i1, b1 = blocks[1]
        i2, b2 = blocks[256]
        np.testing.assert_array_equal(b1[:, i2], b2[:, i1].T)
