"""ctypes access to the CPU oracle (oracle/_build/liboracle.so, int32) and to
the reference's own CPU code (oracle/_ref/libfapref.so, float) -- checkers
only.  Both are built by oracle/Makefile (``make -C oracle``)."""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfapref.so")
INF = 0x3FFFFFFF
REF_INF = np.float32(1e8)

_pi = C.POINTER(C.c_int32)
_pf = C.POINTER(C.c_float)
_i, _i64 = C.c_int32, C.c_int64


def _p(a):
    return None if a is None else a.ctypes.data_as(_pi)


def _f(a):
    return None if a is None else a.ctypes.data_as(_pf)


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_build/liboracle.so"],
                           check=True, capture_output=True)
        _orc = C.CDLL(ORACLE_SO)
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_maxvalue.restype = C.c_float
    return _ref


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------- int32 oracle ----------------

def dijkstra(src, ro, col, w, with_path=False):
    V = len(ro) - 1
    d = np.empty(V, np.int32)
    p = np.empty(V, np.int32) if with_path else None
    ro, col, w = i32(ro), i32(col), i32(w)
    oracle().orc_dijkstra(_i(src), _i(V), _p(ro), _p(col), _p(w), _p(d), _p(p))
    return (d, p) if with_path else d


def batched_sssp(sources, ro, col, w, nthreads=0):
    V = len(ro) - 1
    s = i32(sources)
    out = np.empty((len(s), V), np.int32)
    ro, col, w = i32(ro), i32(col), i32(w)
    oracle().orc_batched_sssp(_p(s), _i(len(s)), _i(V), _p(ro), _p(col), _p(w), _p(out),
                              _i64(V), None, C.c_int(nthreads))
    return out


def floyd(mat, with_path=None):
    m = np.array(mat, dtype=np.int32, copy=True, order="C")
    n = m.shape[0]
    path = None if with_path is None else np.array(with_path, np.int32, copy=True, order="C")
    oracle().orc_floyd(_i(n), _p(m), _p(path))
    return (m, path) if with_path is not None else m


def min_plus(A, B, B_path=None):
    A, B = i32(A), i32(B)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.empty((M, N), np.int32)
    Cp = np.empty((M, N), np.int32) if B_path is not None else None
    oracle().orc_min_plus(_p(A), _p(B), _p(None if B_path is None else i32(B_path)), _p(Cm),
                          _p(Cp), _i(M), _i(N), _i(K))
    return (Cm, Cp) if B_path is not None else Cm


def min_plus_acc(A, B, Cin):
    A, B = i32(A), i32(B)
    Cm = np.array(Cin, np.int32, copy=True, order="C")
    M, K = A.shape
    N = B.shape[1]
    oracle().orc_min_plus_acc(_p(A), _i64(K), _p(B), _i64(N), _p(Cm), _i64(N), _i(M), _i(N), _i(K))
    return Cm


def find_boundary(K, graph_id, ro, col):
    V = len(ro) - 1
    out = np.empty(V, np.int32)
    oracle().orc_find_boundary(_i(K), _i(V), _p(i32(graph_id)), _p(i32(ro)), _p(i32(col)), _p(out))
    return out


def sort_and_encode(K, graph_id, is_bdy):
    V = len(graph_id)
    arrs = [np.empty(K + 1, np.int32) for _ in range(4)]
    st2ed = np.empty(V, np.int32)
    ed2st = np.empty(V, np.int32)
    oracle().orc_sort_and_encode(_i(K), _i(V), _p(i32(graph_id)), _p(i32(is_bdy)),
                                 *[_p(a) for a in arrs], _p(st2ed), _p(ed2st))
    return dict(ver_num=arrs[0], ver_off=arrs[1], bdy_num=arrs[2], bdy_off=arrs[3],
                st2ed=st2ed, ed2st=ed2st)


def precondition(K, graph_id, ro, col):
    b = find_boundary(K, graph_id, ro, col)
    d = sort_and_encode(K, graph_id, b)
    d["is_bdy"] = b
    return d


def submat_build(rowblock, graph_id, start, n_sub, n_bdy, st2ed, ed2st, ro, col, w):
    V = len(ro) - 1
    sub = np.full((n_sub, n_sub), INF, np.int32)
    rb = i32(rowblock)
    oracle().orc_submat_build(_p(sub), None, _p(rb), None, _p(i32(graph_id)), _i(start),
                              _i(n_sub), _i(n_bdy), _i(V), _p(i32(st2ed)), _p(i32(ed2st)),
                              _p(i32(ro)), _p(i32(col)), _p(i32(w)))
    return sub


def submat_decode(sub, rowblock, start, st2ed):
    rb = np.array(rowblock, np.int32, copy=True, order="C")
    n = sub.shape[0]
    oracle().orc_submat_decode(_p(i32(sub)), None, _p(rb), None, _i(start), _i(n), _i(n),
                               _i(rb.shape[1]), _p(i32(st2ed)))
    return rb


def run_part(pid, ro, col, w, graph_id, pc, nthreads=0, with_path=False):
    V = len(ro) - 1
    n = int(pc["ver_num"][pid])
    out = np.empty((n, V), np.int32)
    path = np.empty((n, V), np.int32) if with_path else None
    oracle().orc_run_part(_i(pid), _i(V), _p(i32(ro)), _p(i32(col)), _p(i32(w)),
                          _p(i32(graph_id)), _p(pc["st2ed"]), _p(pc["ed2st"]), _p(pc["ver_num"]),
                          _p(pc["ver_off"]), _p(pc["bdy_num"]), _p(out), _p(path),
                          C.c_int(nthreads))
    return (out, path) if with_path else out


def check_pred(dist, pred, own, ro, col, w) -> int:
    """Number of cells of predecessor rows `pred` that break the reference's
    path semantics for distance rows `dist` (0 = valid)."""
    D, P = i32(dist), i32(pred)
    o = i32(own)
    f = oracle().orc_check_pred
    f.restype = C.c_int64
    return int(f(_p(D), _i64(D.shape[1]), _p(P), _i64(P.shape[1]), _p(o), _i(len(o)),
                 _i(len(ro) - 1), _p(i32(ro)), _p(i32(col)), _p(i32(w))))


# ---------------- reference CPU code (float) ----------------

def to_ref(a):
    a = np.asarray(a)
    return np.where(a >= INF, REF_INF, a.astype(np.float32)).astype(np.float32)


def from_ref(a):
    a = np.asarray(a, np.float32)
    out = a.astype(np.int64)
    out[a >= REF_INF] = INF
    return out.astype(np.int32)


def ref_read(name: str, directed: bool, weighted: bool):
    r = ref()
    n, m = C.c_int(), C.c_int()
    r.ref_read_ver_edges(name.encode(), int(directed), int(weighted), C.byref(n), C.byref(m))
    adj = np.zeros(n.value, np.int32)
    ro = np.zeros(n.value + 1, np.int32)
    col = np.zeros(m.value, np.int32)
    w = np.zeros(m.value, np.float32)
    r.ref_read_mat_file(name.encode(), int(directed), int(weighted), n, m, _p(adj), _p(ro),
                        _p(col), _f(w))
    return adj, ro, col, w


def ref_dijkstra(src, ro, col, w):
    V = len(ro) - 1
    ro = i32(ro)
    adj = np.diff(ro).astype(np.int32)
    d = np.empty(V, np.float32)
    p = np.empty(V, np.int32)
    ref().ref_dijkstra_path(_i(src), _i(V), _p(adj), _p(ro), _p(i32(col)),
                            _f(to_ref(w)), _f(d), _p(p))
    return d, p


def ref_floyd(mat):
    m = to_ref(mat).copy(order="C")
    n = m.shape[0]
    path = np.zeros((n, n), np.int32)
    ref().ref_floyd_cpu_path(_i(n), _f(m), _p(path))
    return m


def ref_min_plus(A, B):
    A, B = to_ref(A), to_ref(B)
    M, K = A.shape
    N = B.shape[1]
    Cm = np.empty((M, N), np.float32)
    bp = np.zeros((K, N), np.int32)
    cp = np.empty((M, N), np.int32)
    ref().ref_min_plus_cpu_advanced(_f(A), _f(B), _p(bp), _f(Cm), _p(cp), _i(M), _i(N), _i(K))
    return Cm


def ref_min_plus_path(A, B, B_path):
    """The reference's min_plus_path_CPU_advanced with its path matrix."""
    A, B = to_ref(A), to_ref(B)
    M, K = A.shape
    N = B.shape[1]
    Cm = np.empty((M, N), np.float32)
    cp = np.empty((M, N), np.int32)
    ref().ref_min_plus_cpu_advanced(_f(A), _f(B), _p(i32(B_path)), _f(Cm), _p(cp), _i(M), _i(N),
                                    _i(K))
    return Cm, cp


def fw_hop_check(d_in, d_out, hop):
    """Cells of an FW hop matrix (identity labels: hop[i][j] = the m of the
    last hop m -> j, or the input label where the cell was not improved)
    that break the closure's path contract; 0 = valid.  Checked as
    predecessor rows of the input matrix's digraph (arc m -> j of weight
    d_in[m][j] for every finite off-diagonal entry) by the oracle's checker,
    which also walks every hop chain back to its row."""
    d_in = i32(d_in)
    n = d_in.shape[0]
    fin = (d_in < INF) & ~np.eye(n, dtype=bool)
    ro = np.concatenate([[0], np.cumsum(fin.sum(1))]).astype(np.int32)
    col = np.nonzero(fin)[1].astype(np.int32)
    w = d_in[fin].astype(np.int32)
    return check_pred(d_out, hop, np.arange(n, dtype=np.int32), ro, col, w)


def ref_precondition(K, graph_id, ro, col):
    V = len(ro) - 1
    ro = i32(ro)
    adj = np.diff(ro).astype(np.int32)
    w = np.ones(len(col), np.float32)
    out = dict(is_bdy=np.empty(V, np.int32), st2ed=np.empty(V, np.int32),
               ed2st=np.empty(V, np.int32))
    t = [np.zeros(K + 1, np.int32) for _ in range(4)]
    ref().ref_precondition(_i(K), _i(V), _p(i32(graph_id).copy()), _p(adj), _p(ro),
                           _p(i32(col)), _f(w), _p(out["is_bdy"]), *[_p(a) for a in t],
                           _p(out["st2ed"]), _p(out["ed2st"]))
    out.update(ver_num=t[0], ver_off=t[1], bdy_num=t[2], bdy_off=t[3])
    return out


def ref_run_part(pid, ro, col, w, graph_id, pc, with_path=False):
    V = len(ro) - 1
    ro = i32(ro)
    adj = np.diff(ro).astype(np.int32)
    n = int(pc["ver_num"][pid])
    out = np.zeros((n, V), np.float32)
    outp = np.zeros((n, V), np.int32)
    ref().ref_run_part(_i(pid), _i(V), _i(len(col)), _p(adj), _p(ro), _p(i32(col)),
                       _f(to_ref(w)), _p(i32(graph_id).copy()), _p(pc["st2ed"].copy()),
                       _p(pc["ed2st"].copy()), _p(pc["ver_num"].copy()), _p(pc["ver_off"].copy()),
                       _p(pc["bdy_num"].copy()), _f(out), _p(outp))
    return (out, outp) if with_path else out
