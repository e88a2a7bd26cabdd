"""The reference's three stage operators on host float arrays (fap_ops.h).

Same argument meaning as fap::handle_boundry_path (fap/kernel/batch_sssp.h:
254-277), fap::floyd_path (fap/kernel/floyd.h:150-158) and
fap::min_plus_path_advanced (fap/kernel/minplus.h:78-89): float32 arrays,
INF = 1e8, row-major without leading dimensions.  Each call runs on the GPU.
"""

from __future__ import annotations

import numpy as np

from . import check, load, _p32, _pf


def handle_boundry_path(sub_graph: np.ndarray, vertexs: int, edges: int, bdy_num: int,
                        row_offset, col_val, weights, st2ed, offset: int):
    """Fill rows 0..bdy_num-1 of sub_graph (float32, stride vertexs) in place."""
    assert sub_graph.dtype == np.float32 and sub_graph.flags.c_contiguous
    ro = np.ascontiguousarray(row_offset, np.int32)
    c = np.ascontiguousarray(col_val, np.int32)
    w = np.ascontiguousarray(weights, np.float32)
    s = np.ascontiguousarray(st2ed, np.int32)
    check(load().fap_handle_boundry_path(_pf(sub_graph), None, vertexs, edges, bdy_num, None,
                                         _p32(ro), _p32(c), _pf(w), _p32(s), offset),
          "handle_boundry_path")


def floyd_path(num: int, mat: np.ndarray):
    """In-place closure of the num x num float32 matrix."""
    assert mat.dtype == np.float32 and mat.flags.c_contiguous and mat.size == num * num
    check(load().fap_floyd_path(num, _pf(mat), None), "floyd_path")


def min_plus_path_advanced(mat1: np.ndarray, mat2: np.ndarray, res: np.ndarray, M: int, N: int,
                           K: int):
    """res (M x N) = min(INF, mat1 (M x K) (x) mat2 (K x N)), float32, in place."""
    for a in (mat1, mat2, res):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    check(load().fap_min_plus_path_advanced(_pf(mat1), _pf(mat2), None, _pf(res), None, M, N, K),
          "min_plus_path_advanced")
