"""fastapsp_amd -- MI355X (gfx950) subgraph-solve path of FastAPSP.

Python access to libfap_amd.so through ctypes: the C ABI of include/fap_dev.h
(device-pointer kernels), include/fap_graph.h (the fap::fapGraph drop-in) and
include/fap_ops.h (the reference's host float operators).  numpy arrays are
the host format; device memory is managed by the library itself
(``Context.alloc``), so no other GPU runtime is involved.

The library must have been built (``python fastapsp_amd/build.py`` or
``__graft_entry__.build()``); importing the compute API without it raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

INF = 0x3FFFFFFF
REF_INF = 1e8

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfap_amd.so")

_i32, _i64, _u64, _int = C.c_int32, C.c_int64, C.c_uint64, C.c_int
_vp, _sz = C.c_void_p, C.c_size_t
_pi32 = C.POINTER(C.c_int32)
_pf32 = C.POINTER(C.c_float)
_pu64 = C.POINTER(C.c_uint64)



class ServeOpts(C.Structure):
    """fap_serve_opts (include/fap_serve.h)."""
    _fields_ = [("hbm_budget", C.c_int64), ("stage_bytes", C.c_int64), ("lanes", C.c_int32),
                ("max_batch", C.c_int32), ("max_sources", C.c_int32), ("poll_us", C.c_int32),
                ("query_parts", C.c_int32)]


class ServeInfo(C.Structure):
    """fap_serve_info (include/fap_serve.h)."""
    _fields_ = [("state", C.c_int32), ("kind", C.c_int32), ("n", C.c_int32),
                ("wave_size", C.c_int32), ("wave", C.c_int64), ("ld", C.c_int64),
                ("submit_us", C.c_double), ("admit_us", C.c_double), ("done_us", C.c_double)]


class ServeStats(C.Structure):
    """fap_serve_stats (include/fap_serve.h)."""
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "pending", "running",
                                         "waves", "arena_bytes", "arena_used", "arena_peak",
                                         "fw_ops", "sssp_sources")] + \
               [("max_wave", C.c_int32), ("lanes", C.c_int32), ("index_bytes", C.c_int64),
                ("index_build_ms", C.c_double)]


# name -> (restype, argtypes)
_SIGS = {
    "fap_last_error": (C.c_char_p, []),
    "fap_version": (C.c_char_p, []),
    "fap_device_count": (_int, [C.POINTER(_int)]),
    "fap_ctx_create": (_int, [_int, C.POINTER(_vp)]),
    "fap_ctx_destroy": (_int, [_vp]),
    "fap_ctx_stream": (_int, [_vp, C.POINTER(_vp)]),
    "fap_ctx_set_option": (_int, [_vp, C.c_char_p, _i64]),
    "fap_ctx_get_option": (_int, [_vp, C.c_char_p, C.POINTER(_i64)]),
    "fap_ctx_sync": (_int, [_vp]),
    "fap_dev_alloc": (_int, [_vp, _sz, C.POINTER(_vp)]),
    "fap_dev_free": (_int, [_vp, _vp]),
    "fap_memcpy_h2d": (_int, [_vp, _vp, _vp, _sz]),
    "fap_memcpy_d2h": (_int, [_vp, _vp, _vp, _sz]),
    "fap_memcpy_d2d": (_int, [_vp, _vp, _vp, _sz]),
    "fap_fill_i32": (_int, [_vp, _vp, _i64, _i32, _vp]),
    "fap_copy_i32": (_int, [_vp, _vp, _vp, _i64, _vp]),
    "fap_checksum_i32": (_int, [_vp, _vp, _i64, _pu64, _pu64, _vp]),
    "fap_stream_create": (_int, [_vp, C.POINTER(_vp)]),
    "fap_stream_destroy": (_int, [_vp, _vp]),
    "fap_stream_sync": (_int, [_vp]),
    "fap_event_create": (_int, [C.POINTER(_vp)]),
    "fap_event_destroy": (_int, [_vp]),
    "fap_event_record": (_int, [_vp, _vp]),
    "fap_event_elapsed": (_int, [_vp, _vp, C.POINTER(C.c_float)]),
    "fap_graph_upload": (_int, [_vp, _i32, _i64, _pi32, _pi32, _pi32, C.POINTER(_vp)]),
    "fap_graph_set_partition": (_int, [_vp, _vp, _pi32, _pi32, _pi32]),
    "fap_graph_free": (_int, [_vp, _vp]),
    "fap_graph_set_delta": (_int, [_vp, _i32]),
    "fap_graph_info": (_int, [_vp, _pi32, C.POINTER(_i64)]),
    "fap_sssp_multi": (_int, [_vp, _vp, _vp, _i32, _vp, _i64, _vp]),
    "fap_fw_blocked": (_int, [_vp, _vp, _i32, _i64, _vp]),
    "fap_fw_batched": (_int, [_vp, _i32, C.POINTER(_vp), _pi32, C.POINTER(_i64), _vp]),
    "fap_fw_batched_noalloc": (_int, [_vp, _i32, C.POINTER(_vp), _pi32, C.POINTER(_i64), _vp]),
    "fap_fw_batched_persistent": (_int, [_vp, _i32, C.POINTER(_vp), _pi32, C.POINTER(_i64),
                                         _vp]),
    "fap_fw_plan_create": (_int, [_vp, _i32, C.POINTER(_vp), _pi32, C.POINTER(_i64), _int,
                                  C.POINTER(_vp)]),
    "fap_fw_plan_run": (_int, [_vp, _vp]),
    "fap_fw_plan_create_persistent": (_int, [_vp, _i32, C.POINTER(_vp), _pi32, C.POINTER(_i64),
                                             _int, C.POINTER(_vp)]),
    "fap_fw_plan_check": (_int, [_vp]),
    "fap_fw_plan_info": (_int, [_vp, _pi32, C.POINTER(_i64), C.POINTER(_i64)]),
    "fap_fw_plan_destroy": (_int, [_vp]),
    "fap_dense_from_graphs": (_int, [_vp, _i32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_i64),
                                     _vp]),
    "fap_minplus": (_int, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _i32, _i32, _i32, _int, _vp]),
    "fap_gather_submat": (_int, [_vp, _vp, _i32, _i32, _i32, _vp, _i64, _vp, _i64, _vp]),
    "fap_scatter_submat": (_int, [_vp, _vp, _i32, _i32, _i32, _vp, _i64, _vp, _i64, _vp]),
    "fap_subgraph_solve": (_int, [_vp, _vp, _i32, _i32, _i32, _vp, _i64, _vp]),
    "fap_subgraph_solve_batch": (_int, [_vp, _vp, _i32, _pi32, _pi32, _pi32, C.POINTER(_vp),
                                        _i64, _vp]),
    # fap_ops.h
    "fap_handle_boundry_path": (_int, [_pf32, _pi32, _int, _int, _int, _pi32, _pi32, _pi32,
                                       _pf32, _pi32, _int]),
    "fap_floyd_path": (_int, [_int, _pf32, _pi32]),
    "fap_min_plus_path_advanced": (_int, [_pf32, _pf32, _pi32, _pf32, _pi32, _int, _int, _int]),
    # fap_graph.h
    "fapg_create_from_file": (_int, [C.c_char_p, _int, _int, _i32, C.POINTER(_vp)]),
    "fapg_create_from_csr": (_int, [_i32, _i64, _pi32, _pi32, _pf32, _i32, C.POINTER(_vp)]),
    "fapg_destroy": (_int, [_vp]),
    "fapg_set_partition": (_int, [_vp, _pi32]),
    "fapg_precondition": (_int, [_vp]),
    "fapg_is_split": (_int, [_vp, C.POINTER(_int)]),
    "fapg_solve": (_int, [_vp, _int]),
    "fapg_solve_subgraph": (_int, [_vp, _i32, _int]),
    "fapg_set_batched": (_int, [_vp, _int]),
    "fapg_solve_subgraphs": (_int, [_vp, _i32, _pi32, _int]),
    "fapg_num_vertices": (_int, [_vp, _pi32]),
    "fapg_num_edges": (_int, [_vp, C.POINTER(_i64)]),
    "fapg_num_parts": (_int, [_vp, _pi32]),
    "fapg_is_directed": (_int, [_vp, C.POINTER(_int)]),
    "fapg_used_metis": (_int, [_vp, C.POINTER(_int)]),
    "fapg_get_csr": (_int, [_vp, _pi32, _pi32, _pi32]),
    "fapg_get_graph_id": (_int, [_vp, _pi32]),
    "fapg_get_boundary_flags": (_int, [_vp, _pi32]),
    "fapg_get_mapping": (_int, [_vp, _pi32, _pi32]),
    "fapg_get_part_info": (_int, [_vp, _pi32, _pi32, _pi32, _pi32]),
    "fapg_get_current_subgraph_id": (_int, [_vp, _pi32]),
    "fapg_get_subgraph_index": (_int, [_vp, _i32, _pi32]),
    "fapg_get_subgraph_distance_i32": (_int, [_vp, _i32, _pi32]),
    "fapg_get_subgraph_distance_f32": (_int, [_vp, _i32, _pf32]),
    "fapg_get_all_distance_i32": (_int, [_vp, _pi32]),
    "fapg_get_all_distance_f32": (_int, [_vp, _pf32]),
    "fapg_get_distance_from_one_point": (_int, [_vp, _i32, _pf32]),
    "fapg_distance_p2p": (_int, [_vp, _i32, _i32, _pi32]),
    "fapg_get_subgraph_rows_i32": (_int, [_vp, _i32, _i32, _i32, _pi32]),
    "fapg_subgraph_checksum": (_int, [_vp, _i32, _pu64, _pu64]),
    "fapg_get_subgraph_path": (_int, [_vp, _i32, _pi32]),
    "fapg_get_all_path": (_int, [_vp, _pi32]),
    "fapg_get_path_from_one_point": (_int, [_vp, _i32, _pi32]),
    "fapg_path_p2p": (_int, [_vp, _i32, _i32, _pi32, _i32, _pi32]),
    "fap_pred_rows": (_int, [_vp, _vp, _vp, _i64, _vp, _i32, _vp, _i64, _vp]),
    "fap_minplus_path": (_int, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                _i32, _i32, _i32, _vp]),
    "fap_fw_blocked_path": (_int, [_vp, _vp, _i32, _i64, _vp, _i64, _vp]),
    "fap_graph_build_index": (_int, [_vp, _vp, C.POINTER(_i32)]),
    "fap_graph_index_info": (_int, [_vp, C.POINTER(_i32), C.POINTER(_i64)]),
    "fap_subgraph_solve_groups": (_int, [_vp, _vp, _i32, _pi32, _pi32, _pi32, _vp, _i64, _i32]),
    "fap_subgraph_boundary_rows": (_int, [_vp, _vp, _i32, _pi32, _pi32, _pi32, C.POINTER(_vp),
                                          _i64, _vp]),
    "fap_subgraph_pred": (_int, [_vp, _vp, _i32, _i32, _vp, _i64, _vp, _i64, _vp]),
    "fapg_check_result": (_int, [_vp, _i32, C.POINTER(_int)]),
    # serving (include/fap_serve.h)
    "fap_serve_default_opts": (None, [C.POINTER(ServeOpts)]),
    "fap_serve_create": (_int, [_int, C.POINTER(ServeOpts), C.POINTER(_vp)]),
    "fap_serve_destroy": (_int, [_vp]),
    "fap_serve_attach_graph": (_int, [_vp, _i32, _i64, _pi32, _pi32, _pi32]),
    "fap_serve_submit_graph": (_int, [_vp, _i32, _i64, _pi32, _pi32, _pi32,
                                      C.POINTER(_i64)]),
    "fap_serve_submit_queries": (_int, [_vp, _i32, _pi32, _pi32, C.POINTER(_i64)]),
    "fap_serve_wait": (_int, [_vp, _i64, _i32, C.POINTER(ServeInfo)]),
    "fap_serve_answers": (_int, [_vp, _i64, _pi32]),
    "fap_serve_rows": (_int, [_vp, _i64, _i32, _i32, _pi32]),
    "fap_serve_checksum": (_int, [_vp, _i64, _pu64, _pu64]),
    "fap_serve_block": (_int, [_vp, _i64, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "fap_serve_release": (_int, [_vp, _i64]),
    "fap_serve_get_stats": (_int, [_vp, C.POINTER(ServeStats)]),
    # autotuner (include/fap_tune.h)
    "fap_tune_cache_create": (_int, [C.POINTER(_vp)]),
    "fap_tune_cache_destroy": (_int, [_vp]),
    "fap_tune_cache_load": (_int, [_vp, C.c_char_p]),
    "fap_tune_cache_save": (_int, [_vp, C.c_char_p]),
    "fap_tune_cache_size": (_int, [_vp, _pi32]),
    "fap_tune_cache_set": (_int, [_vp, C.c_char_p, _i64, _pi32, _i32, C.c_double]),
    "fap_tune_cache_get": (_int, [_vp, C.c_char_p, _i64, _pi32, _i32,
                                  C.POINTER(C.c_double)]),
    "fap_tune_key_fw": (_i64, [_i32]),
    "fap_tune_key_sssp": (_i64, [_i32, _i64, _i32]),
    "fap_tune_key_batch": (_i64, [_i32, _pi32]),
    "fap_ctx_set_tune_cache": (_int, [_vp, _vp]),
    "fap_tune_fw": (_int, [_vp, _i32, _i32]),
    "fap_tune_sssp": (_int, [_vp, _vp, _pi32, _i32, _i32]),
    "fap_tune_batch": (_int, [_vp, _i32, _pi32, _i32]),
    "fap_probe_minplus_ceiling": (_int, [_vp, C.POINTER(C.c_double)]),
}

_lib = None


class FapError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: str | None = None):
    """Load libfap_amd.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FapError(-4, f"{p} is missing: build it with `python fastapsp_amd/build.py`")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().fap_last_error().decode()


def check(rc: int, what: str = ""):
    if rc != 0:
        raise FapError(rc, f"{what}: {last_error()}" if what else last_error())


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_pi32)


def _pf(a: np.ndarray):
    return a.ctypes.data_as(_pf32)


class DeviceArray:
    """int32 device buffer owned by a Context (shape is host bookkeeping)."""

    def __init__(self, ctx: "Context", shape, ptr: int):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.ptr = ptr

    @property
    def size(self) -> int:
        return int(np.prod(self.shape)) if self.shape else 1

    @property
    def nbytes(self) -> int:
        return 4 * self.size

    def offset(self, elems: int) -> int:
        return self.ptr + 4 * int(elems)

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.int32)
        if out.size:
            check(load().fap_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def set(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=np.int32)
        assert h.size == self.size, (h.shape, self.shape)
        if h.size:
            check(load().fap_memcpy_h2d(self.ctx.handle, self.ptr, h.ctypes.data, h.nbytes), "h2d")

    def free(self):
        if self.ptr:
            load().fap_dev_free(self.ctx.handle, self.ptr)
            self.ptr = 0


class TuneCache:
    """The autotuner's cache (fap_tune_cache): host-side, no device needed."""

    _NP = {"fw": 2, "sssp": 2, "batch": 1}

    def __init__(self, path: str | None = None):
        h = _vp()
        check(load().fap_tune_cache_create(C.byref(h)), "fap_tune_cache_create")
        self.handle = h.value
        if path:
            self.load(path)

    def load(self, path: str):
        check(load().fap_tune_cache_load(self.handle, path.encode()), "tune cache load")

    def save(self, path: str):
        check(load().fap_tune_cache_save(self.handle, path.encode()), "tune cache save")

    def __len__(self):
        n = C.c_int32()
        check(load().fap_tune_cache_size(self.handle, C.byref(n)))
        return n.value

    def set(self, kind: str, key: int, params, ms: float):
        p = np.ascontiguousarray(params, dtype=np.int32)
        check(load().fap_tune_cache_set(self.handle, kind.encode(), int(key), _p32(p), p.size,
                                        float(ms)), "tune cache set")

    def get(self, kind: str, key: int):
        """(params, ms), or None when the bucket has no entry."""
        p = np.zeros(self._NP.get(kind, 0), np.int32)
        ms = C.c_double()
        rc = load().fap_tune_cache_get(self.handle, kind.encode(), int(key), _p32(p), p.size,
                                       C.byref(ms))
        if rc == -4:
            return None
        check(rc, "tune cache get")
        return tuple(int(x) for x in p), ms.value

    @staticmethod
    def key_fw(n: int) -> int:
        return load().fap_tune_key_fw(int(n))

    @staticmethod
    def key_sssp(V: int, E: int, sources: int) -> int:
        return load().fap_tune_key_sssp(int(V), int(E), int(sources))

    @staticmethod
    def key_batch(sizes) -> int:
        n = np.ascontiguousarray(sizes, dtype=np.int32)
        return load().fap_tune_key_batch(n.size, _p32(n))

    def close(self):
        if self.handle:
            load().fap_tune_cache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A device context (fap_ctx): stream, workspaces, device memory."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _vp()
        check(lib.fap_ctx_create(device, C.byref(h)), "fap_ctx_create")
        self.handle = h.value
        s = _vp()
        check(lib.fap_ctx_stream(self.handle, C.byref(s)), "fap_ctx_stream")
        self.stream = s.value

    def close(self):
        if self.handle:
            load().fap_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(load().fap_ctx_sync(self.handle), "sync")

    def set_option(self, name: str, value: int):
        check(load().fap_ctx_set_option(self.handle, name.encode(), int(value)), "set_option")

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        check(load().fap_ctx_get_option(self.handle, name.encode(), C.byref(v)), "get_option")
        return v.value

    def set_tune_cache(self, cache: "TuneCache | None"):
        """Consult `cache` for every option left at auto (None detaches)."""
        check(load().fap_ctx_set_tune_cache(self.handle,
                                          cache.handle if cache is not None else None),
              "set_tune_cache")
        self._tune = cache  # keep it alive while attached

    def tune_fw(self, n: int, reps: int = 2):
        check(load().fap_tune_fw(self.handle, int(n), int(reps)), "tune_fw")

    def tune_sssp(self, graph: "DeviceGraph", sources, reps: int = 1):
        s = np.ascontiguousarray(sources, dtype=np.int32)
        check(load().fap_tune_sssp(self.handle, graph.handle, _p32(s), s.size, int(reps)),
              "tune_sssp")

    def tune_batch(self, sizes, reps: int = 2):
        n = np.ascontiguousarray(sizes, dtype=np.int32)
        check(load().fap_tune_batch(self.handle, n.size, _p32(n), int(reps)), "tune_batch")

    def probe_minplus_ceiling(self):
        """Measured TOP/s of the min-plus instruction mix (2 ops per update)
        on this device: the VALU ceiling the min-plus and FW kernels run
        against."""
        v = C.c_double()
        check(load().fap_probe_minplus_ceiling(self.handle, C.byref(v)), "probe_minplus_ceiling")
        return v.value

    def alloc(self, shape) -> DeviceArray:
        shape = (shape,) if isinstance(shape, (int, np.integer)) else tuple(shape)
        n = int(np.prod(shape)) if shape else 1
        p = _vp()
        check(load().fap_dev_alloc(self.handle, max(4, 4 * n), C.byref(p)), "alloc")
        return DeviceArray(self, shape, p.value)

    def to_device(self, host: np.ndarray) -> DeviceArray:
        h = np.ascontiguousarray(host, dtype=np.int32)
        d = self.alloc(h.shape)
        d.set(h)
        return d

    def fill(self, d: DeviceArray, value: int, n: int | None = None, ptr: int | None = None):
        check(load().fap_fill_i32(self.handle, ptr or d.ptr, d.size if n is None else n,
                                  int(value), None), "fill")

    def checksum(self, ptr: int, n: int):
        s, w = C.c_uint64(), C.c_uint64()
        check(load().fap_checksum_i32(self.handle, ptr, n, C.byref(s), C.byref(w), None), "checksum")
        return s.value, w.value

    # ---- kernels (device pointers; see fap_dev.h) ----
    def minplus(self, A: int, lda: int, B: int, ldb: int, C_: int, ldc: int, M: int, N: int,
                K: int, accumulate: bool = False, stream=None):
        check(load().fap_minplus(self.handle, A, lda, B, ldb, C_, ldc, M, N, K,
                                 1 if accumulate else 0, stream), "fap_minplus")

    def fw(self, ptr: int, n: int, ld: int, stream=None):
        check(load().fap_fw_blocked(self.handle, ptr, n, ld, stream), "fap_fw_blocked")

    def minplus_path(self, A: int, lda: int, B: int, ldb: int, Bp: int, ldbp: int, C_: int,
                     ldc: int, Cp: int, ldcp: int, M: int, N: int, K: int, stream=None):
        check(load().fap_minplus_path(self.handle, A, lda, B, ldb, Bp, ldbp, C_, ldc, Cp, ldcp,
                                      M, N, K, stream), "fap_minplus_path")

    def fw_path(self, ptr: int, n: int, ld: int, path: int, ldp: int, stream=None):
        check(load().fap_fw_blocked_path(self.handle, ptr, n, ld, path, ldp, stream),
              "fap_fw_blocked_path")

    def fw_batched(self, ptrs, ns, lds, stream=None):
        k = len(ptrs)
        P = (_vp * k)(*ptrs)
        N = np.ascontiguousarray(ns, dtype=np.int32)
        L = np.ascontiguousarray(lds, dtype=np.int64)
        check(load().fap_fw_batched(self.handle, k, P, _p32(N), L.ctypes.data_as(C.POINTER(_i64)),
                                    stream), "fap_fw_batched")

    def fw_batched_persistent(self, ptrs, ns, lds, stream=None):
        """The batched closure as one launch (fap_fw_batched_persistent)."""
        k = len(ptrs)
        P = (_vp * k)(*ptrs)
        N = np.ascontiguousarray(ns, dtype=np.int32)
        L = np.ascontiguousarray(lds, dtype=np.int64)
        check(load().fap_fw_batched_persistent(self.handle, k, P, _p32(N),
                                               L.ctypes.data_as(C.POINTER(_i64)), stream),
              "fap_fw_batched_persistent")

    def fw_plan(self, ptrs, ns, lds, use_graph=True, persistent=False):
        """A reusable batched closure: phase launches (fap_fw_plan_create) or,
        with persistent, one persistent launch per run
        (fap_fw_plan_create_persistent)."""
        k = len(ptrs)
        P = (_vp * k)(*ptrs)
        N = np.ascontiguousarray(ns, dtype=np.int32)
        L = np.ascontiguousarray(lds, dtype=np.int64)
        h = _vp()
        fn = "fap_fw_plan_create_persistent" if persistent else "fap_fw_plan_create"
        check(getattr(load(), fn)(self.handle, k, P, _p32(N), L.ctypes.data_as(C.POINTER(_i64)),
                                  int(use_graph), C.byref(h)), fn)
        return FwPlan(h.value)

    def dense_from_graphs(self, graphs, ptrs, lds, stream=None):
        k = len(graphs)
        G = (_vp * k)(*[g.handle for g in graphs])
        P = (_vp * k)(*ptrs)
        L = np.ascontiguousarray(lds, dtype=np.int64)
        check(load().fap_dense_from_graphs(self.handle, k, G, P, L.ctypes.data_as(C.POINTER(_i64)),
                                           stream), "fap_dense_from_graphs")

    def upload_graph(self, row_offset, col, w) -> "DeviceGraph":
        return DeviceGraph(self, row_offset, col, w)


class FwPlan:
    """A reusable batched closure (fap_fw_plan): run() replays it."""

    def __init__(self, handle):
        self.handle = handle

    def run(self, stream=None):
        check(load().fap_fw_plan_run(self.handle, stream), "fap_fw_plan_run")

    def check(self):
        """Waits for the last run; raises if a persistent run gave up."""
        check(load().fap_fw_plan_check(self.handle), "fap_fw_plan_check")

    def info(self):
        a, b, c = C.c_int32(), C.c_int64(), C.c_int64()
        check(load().fap_fw_plan_info(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return {"launches": a.value, "tasks": b.value, "ops": c.value}

    def close(self):
        if self.handle:
            load().fap_fw_plan_destroy(self.handle)
            self.handle = None


class DeviceGraph:
    """A CSR resident in HBM (fap_graph_dev)."""

    def __init__(self, ctx: Context, row_offset, col, w):
        self.ctx = ctx
        ro = np.ascontiguousarray(row_offset, dtype=np.int32)
        c = np.ascontiguousarray(col, dtype=np.int32)
        ww = np.ascontiguousarray(w, dtype=np.int32)
        self.V = ro.size - 1
        self.E = c.size
        h = _vp()
        check(load().fap_graph_upload(ctx.handle, self.V, self.E, _p32(ro), _p32(c), _p32(ww),
                                      C.byref(h)), "fap_graph_upload")
        self.handle = h.value

    def set_partition(self, graph_id, st2ed, ed2st):
        g = np.ascontiguousarray(graph_id, dtype=np.int32)
        s = np.ascontiguousarray(st2ed, dtype=np.int32)
        e = np.ascontiguousarray(ed2st, dtype=np.int32)
        check(load().fap_graph_set_partition(self.ctx.handle, self.handle, _p32(g), _p32(s),
                                             _p32(e)), "fap_graph_set_partition")

    def set_delta(self, delta: int):
        check(load().fap_graph_set_delta(self.handle, int(delta)), "fap_graph_set_delta")

    def sssp(self, d_sources: int, n_src: int, d_out: int, ld_out: int, stream=None):
        check(load().fap_sssp_multi(self.ctx.handle, self.handle, d_sources, n_src, d_out,
                                    ld_out, stream), "fap_sssp_multi")

    def gather(self, start, n_sub, n_bdy, rowblock, ld_rb, sub, ld_sub, stream=None):
        check(load().fap_gather_submat(self.ctx.handle, self.handle, start, n_sub, n_bdy,
                                       rowblock, ld_rb, sub, ld_sub, stream), "gather")

    def scatter(self, start, rows, n_sub, sub, ld_sub, rowblock, ld_rb, stream=None):
        check(load().fap_scatter_submat(self.ctx.handle, self.handle, start, rows, n_sub, sub,
                                        ld_sub, rowblock, ld_rb, stream), "scatter")

    def subgraph_solve(self, start, n_sub, n_bdy, rowblock, ld_rb, stream=None):
        check(load().fap_subgraph_solve(self.ctx.handle, self.handle, start, n_sub, n_bdy,
                                        rowblock, ld_rb, stream), "fap_subgraph_solve")

    def build_index(self) -> bool:
        b = _i32()
        check(load().fap_graph_build_index(self.ctx.handle, self.handle, C.byref(b)),
              "fap_graph_build_index")
        return bool(b.value)

    def index_info(self):
        nb, by = _i32(), _i64()
        check(load().fap_graph_index_info(self.handle, C.byref(nb), C.byref(by)))
        return nb.value, by.value

    def subgraph_solve_batch(self, starts, n_subs, n_bdys, rowblocks, ld_rb, stream=None):
        k = len(starts)
        S = np.ascontiguousarray(starts, dtype=np.int32)
        N = np.ascontiguousarray(n_subs, dtype=np.int32)
        B = np.ascontiguousarray(n_bdys, dtype=np.int32)
        R = (_vp * k)(*rowblocks)
        check(load().fap_subgraph_solve_batch(self.ctx.handle, self.handle, k, _p32(S), _p32(N),
                                              _p32(B), R, ld_rb, stream),
              "fap_subgraph_solve_batch")

    def boundary_rows(self, starts, n_subs, n_bdys, rowblocks, ld_rb, stream=None):
        """Stage 1 alone: rows 0..b-1 of each part's row block."""
        k = len(starts)
        S = np.ascontiguousarray(starts, dtype=np.int32)
        N = np.ascontiguousarray(n_subs, dtype=np.int32)
        B = np.ascontiguousarray(n_bdys, dtype=np.int32)
        R = (_vp * k)(*rowblocks)
        check(load().fap_subgraph_boundary_rows(self.ctx.handle, self.handle, k, _p32(S), _p32(N),
                                                _p32(B), R, ld_rb, stream),
              "fap_subgraph_boundary_rows")

    def subgraph_solve_groups(self, starts, n_subs, n_bdys, rowblocks, ld_rb, group_parts):
        k = len(starts)
        S = np.ascontiguousarray(starts, dtype=np.int32)
        N = np.ascontiguousarray(n_subs, dtype=np.int32)
        B = np.ascontiguousarray(n_bdys, dtype=np.int32)
        R = (_vp * k)(*rowblocks)
        check(load().fap_subgraph_solve_groups(self.ctx.handle, self.handle, k, _p32(S), _p32(N),
                                               _p32(B), R, ld_rb, int(group_parts)),
              "fap_subgraph_solve_groups")

    def pred_rows(self, d_dist, ld, d_row_vertex, nrows, d_pred, ld_pred, stream=None):
        check(load().fap_pred_rows(self.ctx.handle, self.handle, d_dist, ld, d_row_vertex, nrows,
                                   d_pred, ld_pred, stream), "fap_pred_rows")

    def subgraph_pred(self, start, n_sub, d_rows, ld, d_pred, ld_pred, stream=None):
        check(load().fap_subgraph_pred(self.ctx.handle, self.handle, start, n_sub, d_rows, ld,
                                       d_pred, ld_pred, stream), "fap_subgraph_pred")

    def close(self):
        if self.handle:
            load().fap_graph_free(self.ctx.handle, self.handle)
            self.handle = None


from .graph import Graph  # noqa: E402  (the fap::fapGraph drop-in)
from .serve import Server  # noqa: E402  (request loop)
from . import ops  # noqa: E402,F401
