"""The reference's three stage operators on host float arrays (fap_ops.h).

Same argument meaning as fap::handle_boundry_path (fap/kernel/batch_sssp.h:
254-277), fap::floyd_path (fap/kernel/floyd.h:150-158) and
fap::min_plus_path_advanced (fap/kernel/minplus.h:78-89): float32 arrays,
INF = 1e8, row-major without leading dimensions.  Each call runs on the GPU.
Path arrays (int32) are optional; when given they are written with the
reference's meaning (Dijkstra parents, FW hop labels path[k][j], min-plus
labels mat2_path[k][j] of the first minimising k).
"""

from __future__ import annotations

import numpy as np

from . import check, load, _p32, _pf


def _opt32(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return _p32(a)


def handle_boundry_path(sub_graph: np.ndarray, vertexs: int, edges: int, bdy_num: int,
                        row_offset, col_val, weights, st2ed, offset: int,
                        sub_graph_path: np.ndarray | None = None):
    """Fill rows 0..bdy_num-1 of sub_graph (float32, stride vertexs) in place,
    and of sub_graph_path (int32) with the parents when given."""
    assert sub_graph.dtype == np.float32 and sub_graph.flags.c_contiguous
    ro = np.ascontiguousarray(row_offset, np.int32)
    c = np.ascontiguousarray(col_val, np.int32)
    w = np.ascontiguousarray(weights, np.float32)
    s = np.ascontiguousarray(st2ed, np.int32)
    check(load().fap_handle_boundry_path(_pf(sub_graph), _opt32(sub_graph_path), vertexs, edges,
                                         bdy_num, None, _p32(ro), _p32(c), _pf(w), _p32(s),
                                         offset),
          "handle_boundry_path")


def floyd_path(num: int, mat: np.ndarray, path: np.ndarray | None = None):
    """In-place closure of the num x num float32 matrix (and its int32 path
    labels when given)."""
    assert mat.dtype == np.float32 and mat.flags.c_contiguous and mat.size == num * num
    assert path is None or path.size == num * num
    check(load().fap_floyd_path(num, _pf(mat), _opt32(path)), "floyd_path")


def min_plus_path_advanced(mat1: np.ndarray, mat2: np.ndarray, res: np.ndarray, M: int, N: int,
                           K: int, mat2_path: np.ndarray | None = None,
                           path: np.ndarray | None = None):
    """res (M x N) = min(INF, mat1 (M x K) (x) mat2 (K x N)), float32, in place;
    path (M x N) = mat2_path[k][j] of the first minimising k when given."""
    for a in (mat1, mat2, res):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    check(load().fap_min_plus_path_advanced(_pf(mat1), _pf(mat2), _opt32(mat2_path), _pf(res),
                                            _opt32(path), M, N, K),
          "min_plus_path_advanced")
