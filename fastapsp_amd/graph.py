"""Graph -- Python handle on the C++ fap::fapGraph drop-in (fap_graph.h).

Method names follow the reference's fap/fap.h:79-124 (preCondition, isSplit,
solve, solveSubGraph, getMapping, getSubGraphIndex, getSubGraphDistance, ...)
so that tests read like the reference's example program
(app/singleNodeExample.cpp:40-68).  Arrays are numpy.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import FapError, check, load, _p32, _pf, _vp


class Graph:
    def __init__(self, handle: int):
        self.handle = handle

    # ---- construction (fap/fap.cpp:27-65) ----
    @classmethod
    def from_file(cls, name: str, directed: bool, weighted: bool, K: int) -> "Graph":
        h = _vp()
        check(load().fapg_create_from_file(name.encode(), int(directed), int(weighted), int(K),
                                           C.byref(h)), "fapGraph(file)")
        return cls(h.value)

    @classmethod
    def from_csr(cls, row_offset, col, weight, K: int) -> "Graph":
        ro = np.ascontiguousarray(row_offset, dtype=np.int32)
        c = np.ascontiguousarray(col, dtype=np.int32)
        w = np.ascontiguousarray(weight, dtype=np.float32)
        h = _vp()
        check(load().fapg_create_from_csr(ro.size - 1, c.size, _p32(ro), _p32(c), _pf(w), int(K),
                                          C.byref(h)), "fapGraph(csr)")
        return cls(h.value)

    def close(self):
        if self.handle:
            load().fapg_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- sizes ----
    @property
    def num_vertexs(self) -> int:
        v = C.c_int32()
        check(load().fapg_num_vertices(self.handle, C.byref(v)))
        return v.value

    @property
    def num_edges(self) -> int:
        v = C.c_int64()
        check(load().fapg_num_edges(self.handle, C.byref(v)))
        return v.value

    @property
    def K(self) -> int:
        v = C.c_int32()
        check(load().fapg_num_parts(self.handle, C.byref(v)))
        return v.value

    @property
    def directed(self) -> bool:
        v = C.c_int()
        check(load().fapg_is_directed(self.handle, C.byref(v)))
        return bool(v.value)

    def used_metis(self) -> bool:
        v = C.c_int()
        check(load().fapg_used_metis(self.handle, C.byref(v)))
        return bool(v.value)

    def csr(self):
        V, E = self.num_vertexs, self.num_edges
        ro = np.empty(V + 1, np.int32)
        col = np.empty(E, np.int32)
        w = np.empty(E, np.int32)
        check(load().fapg_get_csr(self.handle, _p32(ro), _p32(col), _p32(w)), "csr")
        return ro, col, w

    # ---- reference API ----
    def setPartition(self, graph_id):
        g = np.ascontiguousarray(graph_id, dtype=np.int32)
        check(load().fapg_set_partition(self.handle, _p32(g)), "setPartition")

    def preCondition(self):
        check(load().fapg_precondition(self.handle), "preCondition")

    def isSplit(self) -> bool:
        v = C.c_int()
        check(load().fapg_is_split(self.handle, C.byref(v)))
        return bool(v.value)

    def getGraphId(self) -> np.ndarray:
        out = np.empty(self.num_vertexs, np.int32)
        check(load().fapg_get_graph_id(self.handle, _p32(out)), "getGraphId")
        return out

    def boundary_flags(self) -> np.ndarray:
        out = np.empty(self.num_vertexs, np.int32)
        check(load().fapg_get_boundary_flags(self.handle, _p32(out)))
        return out

    def solve(self, is_path_needed: bool = False):
        check(load().fapg_solve(self.handle, int(is_path_needed)), "solve")

    def solveSubGraph(self, sub_graph_id: int, is_path_needed: bool = False):
        check(load().fapg_solve_subgraph(self.handle, int(sub_graph_id), int(is_path_needed)),
              "solveSubGraph")

    def setBatched(self, batched: bool):
        check(load().fapg_set_batched(self.handle, int(batched)))

    def solveSubGraphs(self, ids, is_path_needed: bool = False):
        a = np.ascontiguousarray(ids, dtype=np.int32)
        check(load().fapg_solve_subgraphs(self.handle, a.size, _p32(a), int(is_path_needed)),
              "solveSubGraphs")

    def getMapping(self) -> np.ndarray:
        out = np.empty(self.num_vertexs, np.int32)
        check(load().fapg_get_mapping(self.handle, _p32(out), None), "getMapping")
        return out

    def part_info(self):
        K = self.K
        arrs = [np.empty(K + 1, np.int32) for _ in range(4)]
        check(load().fapg_get_part_info(self.handle, *[_p32(a) for a in arrs]))
        return tuple(arrs)  # ver_num, ver_off, bdy_num, bdy_off

    def getCurrentSubGraphId(self) -> int:
        v = C.c_int32()
        check(load().fapg_get_current_subgraph_id(self.handle, C.byref(v)))
        return v.value

    def getSubGraphIndex(self, sub_graph_id: int) -> np.ndarray:
        vn = self.part_info()[0]
        out = np.empty(int(vn[sub_graph_id]), np.int32)
        check(load().fapg_get_subgraph_index(self.handle, int(sub_graph_id), _p32(out)))
        return out

    def getSubGraphDistanceInt32(self, sub_graph_id: int) -> np.ndarray:
        n = int(self.part_info()[0][sub_graph_id])
        out = np.empty((n, self.num_vertexs), np.int32)
        check(load().fapg_get_subgraph_distance_i32(self.handle, int(sub_graph_id), _p32(out)),
              "getSubGraphDistance")
        return out

    def getSubGraphRowsInt32(self, sub_graph_id: int, row0: int, nrows: int) -> np.ndarray:
        out = np.empty((nrows, self.num_vertexs), np.int32)
        check(load().fapg_get_subgraph_rows_i32(self.handle, int(sub_graph_id), int(row0),
                                                int(nrows), _p32(out)), "getSubGraphRows")
        return out

    def getSubGraphDistance(self, sub_graph_id: int) -> np.ndarray:
        n = int(self.part_info()[0][sub_graph_id])
        out = np.empty((n, self.num_vertexs), np.float32)
        check(load().fapg_get_subgraph_distance_f32(self.handle, int(sub_graph_id), _pf(out)),
              "getSubGraphDistance")
        return out

    def getAllDistanceInt32(self) -> np.ndarray:
        V = self.num_vertexs
        out = np.empty((V, V), np.int32)
        check(load().fapg_get_all_distance_i32(self.handle, _p32(out)), "getAllDistance")
        return out

    def getAllDistance(self) -> np.ndarray:
        V = self.num_vertexs
        out = np.empty((V, V), np.float32)
        check(load().fapg_get_all_distance_f32(self.handle, _pf(out)), "getAllDistance")
        return out

    def getDistanceFromOnePoint(self, v: int) -> np.ndarray:
        out = np.empty(self.num_vertexs, np.float32)
        check(load().fapg_get_distance_from_one_point(self.handle, int(v), _pf(out)))
        return out

    def getDistanceP2P(self, src: int, dst: int) -> int:
        v = C.c_int32()
        check(load().fapg_distance_p2p(self.handle, int(src), int(dst), C.byref(v)))
        return v.value

    # path rows (solved with is_path_needed=True)
    def getSubGraphPath(self, sub_graph_id: int) -> np.ndarray:
        n = int(self.part_info()[0][sub_graph_id])
        out = np.empty((n, self.num_vertexs), np.int32)
        check(load().fapg_get_subgraph_path(self.handle, int(sub_graph_id), _p32(out)),
              "getSubGraphPath")
        return out

    def getAllPath(self) -> np.ndarray:
        V = self.num_vertexs
        out = np.empty((V, V), np.int32)
        check(load().fapg_get_all_path(self.handle, _p32(out)), "getAllPath")
        return out

    def getPathFromOnePoint(self, v: int) -> np.ndarray:
        out = np.empty(self.num_vertexs, np.int32)
        check(load().fapg_get_path_from_one_point(self.handle, int(v), _p32(out)),
              "getPathFromOnePoint")
        return out

    def getPathP2P(self, src: int, dst: int) -> np.ndarray:
        """Vertices src .. dst of a shortest path (empty when unreachable)."""
        cap = self.num_vertexs
        out = np.empty(cap, np.int32)
        n = C.c_int32()
        check(load().fapg_path_p2p(self.handle, int(src), int(dst), _p32(out), cap, C.byref(n)),
              "getPathP2P")
        return out[: n.value].copy()

    def checksum(self, sub_graph_id: int):
        s, w = C.c_uint64(), C.c_uint64()
        check(load().fapg_subgraph_checksum(self.handle, int(sub_graph_id), C.byref(s), C.byref(w)))
        return s.value, w.value

    def check_result(self, sub_graph_id: int) -> bool:
        ok = C.c_int()
        check(load().fapg_check_result(self.handle, int(sub_graph_id), C.byref(ok)), "check_result")
        return bool(ok.value)
