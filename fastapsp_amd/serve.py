"""Request loop on one device (include/fap_serve.h) from Python.

    with Server(hbm_budget=8 << 30) as srv:
        srv.attach_graph(ro, col, w)             # resident graph for point queries
        a = srv.submit_graph(ro2, col2, w2)      # whole-graph APSP
        b = srv.submit_queries([0, 5], [9, 7])   # (src, dst) pairs
        srv.wait(a); srv.wait(b)
        D = srv.rows(a)                          # n x n int32
        d = srv.answers(b)                       # 2 int32
        srv.release(a); srv.release(b)

No counterpart in the reference: it stands in for the multi-task driver
(fap/fap_more_task_mpi.cpp:34-91) on one device and answers the point
queries that fapGraph::getDistanceP2P leaves unimplemented
(fap/fap.cpp:263-265).  Every call goes to the native server in
libfap_amd.so; nothing here computes distances.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import ServeInfo, ServeOpts, ServeStats, _p32, check, load

_STATES = {0: "pending", 1: "running", 2: "done", 3: "failed"}
_KINDS = {0: "graph", 1: "queries"}


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _info(i: ServeInfo) -> dict:
    return {"state": _STATES.get(i.state, i.state), "kind": _KINDS.get(i.kind, i.kind),
            "n": i.n, "wave": i.wave, "wave_size": i.wave_size, "ld": i.ld,
            "submit_us": i.submit_us, "admit_us": i.admit_us, "done_us": i.done_us,
            "latency_us": i.done_us - i.submit_us, "queue_us": i.admit_us - i.submit_us}


class Server:
    def __init__(self, device: int = 0, hbm_budget: int | None = None,
                 stage_bytes: int | None = None, lanes: int | None = None,
                 max_batch: int | None = None, max_sources: int | None = None,
                 poll_us: int | None = None, query_parts: int | None = None):
        lib = load()
        o = ServeOpts()
        lib.fap_serve_default_opts(C.byref(o))
        for k, v in (("hbm_budget", hbm_budget), ("stage_bytes", stage_bytes),
                     ("lanes", lanes), ("max_batch", max_batch),
                     ("max_sources", max_sources), ("poll_us", poll_us),
                     ("query_parts", query_parts)):
            if v is not None:
                setattr(o, k, int(v))
        self.opts = o
        h = C.c_void_p()
        check(lib.fap_serve_create(int(device), C.byref(o), C.byref(h)), "fap_serve_create")
        self.handle = h
        self._n = {}

    # -- inputs -------------------------------------------------------------
    def attach_graph(self, row_offset, col, w):
        ro, c, ww = _i32(row_offset), _i32(col), _i32(w)
        check(load().fap_serve_attach_graph(self.handle, ro.size - 1, c.size, _p32(ro), _p32(c),
                                            _p32(ww)), "fap_serve_attach_graph")

    def submit_graph(self, row_offset, col, w) -> int:
        ro, c, ww = _i32(row_offset), _i32(col), _i32(w)
        rid = C.c_int64()
        check(load().fap_serve_submit_graph(self.handle, ro.size - 1, c.size, _p32(ro), _p32(c),
                                            _p32(ww), C.byref(rid)), "fap_serve_submit_graph")
        self._n[rid.value] = ro.size - 1
        return rid.value

    def submit_queries(self, src, dst) -> int:
        s, d = _i32(src), _i32(dst)
        if s.size != d.size:
            raise ValueError("src and dst differ in length")
        rid = C.c_int64()
        check(load().fap_serve_submit_queries(self.handle, s.size, _p32(s), _p32(d),
                                              C.byref(rid)), "fap_serve_submit_queries")
        self._n[rid.value] = s.size
        return rid.value

    # -- results ------------------------------------------------------------
    def wait(self, rid: int, timeout_ms: int = 120000) -> dict:
        i = ServeInfo()
        check(load().fap_serve_wait(self.handle, rid, int(timeout_ms), C.byref(i)),
              f"request {rid}")
        return _info(i)

    def poll(self, rid: int) -> dict:
        """Non-blocking state of a request (raises if it failed)."""
        i = ServeInfo()
        rc = load().fap_serve_wait(self.handle, rid, 0, C.byref(i))
        if rc not in (0, -7):
            check(rc, f"request {rid}")
        return _info(i)

    def answers(self, rid: int) -> np.ndarray:
        out = np.empty(self._n[rid], np.int32)
        check(load().fap_serve_answers(self.handle, rid, _p32(out)), "fap_serve_answers")
        return out

    def rows(self, rid: int, row0: int = 0, nrows: int | None = None) -> np.ndarray:
        n = self._n[rid]
        nrows = n - row0 if nrows is None else nrows
        out = np.empty((nrows, n), np.int32)
        check(load().fap_serve_rows(self.handle, rid, row0, nrows, _p32(out)), "fap_serve_rows")
        return out

    def checksum(self, rid: int):
        s, ws = C.c_uint64(), C.c_uint64()
        check(load().fap_serve_checksum(self.handle, rid, C.byref(s), C.byref(ws)),
              "fap_serve_checksum")
        return s.value, ws.value

    def block(self, rid: int):
        p, ld, n = C.c_void_p(), C.c_int64(), C.c_int32()
        check(load().fap_serve_block(self.handle, rid, C.byref(p), C.byref(ld), C.byref(n)),
              "fap_serve_block")
        return p.value, ld.value, n.value

    def release(self, rid: int):
        check(load().fap_serve_release(self.handle, rid), "fap_serve_release")
        self._n.pop(rid, None)

    def stats(self) -> dict:
        st = ServeStats()
        check(load().fap_serve_get_stats(self.handle, C.byref(st)), "fap_serve_get_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def close(self):
        if self.handle:
            load().fap_serve_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
