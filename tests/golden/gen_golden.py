"""Writes the golden vectors under tests/golden/ (run once, here; committed).

Expected outputs come from scipy.sparse.csgraph (Dijkstra) and numpy brute
force -- neither the reference nor this repository's code.  Inputs come from
fastapsp_amd.synth (seeded, portable).  The reference holds no golden vectors
of its own for this path; its one fixture graph (luxembourg_osm.mtx, read
only where the FastAPSP checkout is present) contributes two SSSP rows and
is kept itself as its edge lines (luxembourg_osm.npz).

    python tests/golden/gen_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from fastapsp_amd import synth  # noqa: E402

INF = synth.INF
LUX = "/root/reference/app/graph/unweight-undirected/luxembourg_osm.mtx"


def apsp_scipy(V, ro, col, w, directed=True, indices=None):
    m = csr_matrix((w.astype(np.float64), col, ro), shape=(V, V))
    d = shortest_path(m, method="D", directed=directed, indices=indices)
    out = np.where(np.isinf(d), INF, d).astype(np.int64)
    assert out.max() <= INF
    return out.astype(np.int32)


def main():
    # 1. 8x8 grid, 10% of edges removed, weights 1..9, undirected
    u, v, w = synth.grid_graph(8, 8, 0.10, 0.0, 1, 9, 11)
    ro, col, wt = synth.to_csr(64, u, v, w, directed=False)
    np.savez_compressed(os.path.join(HERE, "grid8.npz"), u=u, v=v, w=w, row_offset=ro, col=col,
                        wt=wt, apsp=apsp_scipy(64, ro, col, wt))
    # 2. directed random geometric graph, n = 100, weights 1..255
    au, av, aw = synth.rgg_graph(100, 6, 1, 255, 12)
    ro, col, wt = synth.to_csr(100, au, av, aw, directed=True)
    np.savez_compressed(os.path.join(HERE, "rgg100.npz"), u=au, v=av, w=aw, row_offset=ro,
                        col=col, wt=wt, apsp=apsp_scipy(100, ro, col, wt))
    # 3. dense 70 x 70, arc probability 0.1, weights 1..1000, diagonal 0
    m = synth.dense_matrix(70, 0.1, 1, 1000, 13)
    g = np.where(m >= INF, 0, m).astype(np.float64)   # csgraph: 0 = no arc
    d = shortest_path(csr_matrix(g), method="D", directed=True)
    closure = np.where(np.isinf(d), INF, d).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "dense70.npz"), mat=m, closure=closure)
    # 4. min-plus 13 x 37 (x) 37 x 29 with INF entries, brute force
    ra = synth.stream(14, 1, 13 * 37)
    rb = synth.stream(14, 2, 37 * 29)
    A = synth.uniform_int(ra, 0, 1000).reshape(13, 37)
    B = synth.uniform_int(rb, 0, 1000).reshape(37, 29)
    A[synth.uniform01(synth.stream(14, 3, A.size)).reshape(A.shape) < 0.2] = INF
    B[synth.uniform01(synth.stream(14, 4, B.size)).reshape(B.shape) < 0.2] = INF
    Cb = np.minimum(INF, (A[:, :, None] + B[None, :, :]).min(axis=1))
    np.savez_compressed(os.path.join(HERE, "minplus_13x37x29.npz"), A=A.astype(np.int32),
                        B=B.astype(np.int32), C=Cb.astype(np.int32))
    # 5. two SSSP rows of the reference's fixture graph (unweighted, undirected)
    if os.path.exists(LUX):
        with open(LUX) as f:
            V, _, m_ = map(int, f.readline().split())
            e = np.loadtxt(f, dtype=np.int64).reshape(-1, 2) - 1
        ro, col, wt = synth.to_csr(V, e[:, 0], e[:, 1], np.ones(len(e), np.int64), False)
        src = np.array([0, 57000], dtype=np.int32)
        rows = apsp_scipy(V, ro, col, wt, directed=True, indices=src)
        np.savez_compressed(os.path.join(HERE, "luxembourg_rows.npz"), sources=src, rows=rows,
                            V=V, lines=m_)
        # 6. the fixture graph itself (its header and edge lines, 1-based as in
        # the file), so that the reference's documented run can be replayed
        # where the reference checkout is absent
        np.savez_compressed(os.path.join(HERE, "luxembourg_osm.npz"), V=V, lines=m_,
                            edges=(e + 1).astype(np.int32))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
