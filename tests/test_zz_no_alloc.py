"""fap_fw_batched_noalloc: the batched closure with its descriptors in a
fixed device array instead of a hipMalloc'd buffer, called through the C ABI
and compared bit for bit with fap_fw_batched (and the oracle's FW) on two
calls in a row, the second with a smaller batch, so the second reads a
scratch that still holds the first call's descriptors."""

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 4 << 20   # FAP_NOALLOC_SCRATCH_BYTES


def _args(ptrs, ns, lds):
    k = len(ptrs)
    P = (C.c_void_p * k)(*[int(p) for p in ptrs])
    N = np.ascontiguousarray(ns, dtype=np.int32)
    L = np.ascontiguousarray(lds, dtype=np.int64)
    return k, P, N, L


def _call(name, ctx, ptrs, ns, lds):
    k, P, N, L = _args(ptrs, ns, lds)
    fn = getattr(fa.load(), name)
    return fn(ctx.handle, k, P, N.ctypes.data_as(C.POINTER(C.c_int32)),
              L.ctypes.data_as(C.POINTER(C.c_int64)), None), (P, N, L)


def _mats(sizes, seed):
    return [synth.dense_matrix(int(n), min(1.0, 6.0 / max(int(n), 1)), 1, 255, seed + j)
            for j, n in enumerate(sizes)]


def _upload(ctx, mats):
    lds = [int(m.shape[0]) + 3 for m in mats]
    tot = sum(m.shape[0] * ld for m, ld in zip(mats, lds))
    buf = ctx.alloc(max(tot, 1))
    host = np.full(max(tot, 1), -7, np.int32)
    off, ptrs, offs = 0, [], []
    for m, ld in zip(mats, lds):
        n = m.shape[0]
        blk = host[off:off + n * ld].reshape(n, ld)
        blk[:, :n] = m
        ptrs.append(buf.offset(off))
        offs.append(off)
        off += n * ld
    buf.set(host)
    return buf, ptrs, lds, offs


def _read(buf, mats, lds, offs):
    h = buf.get()
    out = []
    for m, ld, off in zip(mats, lds, offs):
        n = m.shape[0]
        blk = h[off:off + n * ld].reshape(n, ld)
        assert (blk[:, n:] == -7).all()
        out.append(blk[:, :n].copy())
    return out


def test_noalloc_batched_closure_equals_fw_batched(gpu_ctx):
    rng = np.random.default_rng(77)
    first = np.r_[rng.integers(1, 320, 40), [64, 65, 128, 129, 1, 2]]
    second = rng.integers(1, 150, 12)
    for call, sizes in enumerate((first, second)):
        mats = _mats(sizes, 500 + 100 * call)
        a = _upload(gpu_ctx, mats)
        b = _upload(gpu_ctx, mats)
        rc, _keep = _call("fap_fw_batched_noalloc", gpu_ctx, a[1], sizes, a[2])
        assert rc == 0, fa.load().fap_last_error()
        rc, _keep = _call("fap_fw_batched", gpu_ctx, b[1], sizes, b[2])
        assert rc == 0
        got = _read(a[0], mats, a[2], a[3])
        want = _read(b[0], mats, b[2], b[3])
        for j, (g, w, m) in enumerate(zip(got, want, mats)):
            np.testing.assert_array_equal(g, w, err_msg=f"call {call} matrix {j}")
            if m.shape[0] <= 130:
                np.testing.assert_array_equal(g, O.floyd(m), err_msg=f"call {call} matrix {j}")
        a[0].free()
        b[0].free()


def test_noalloc_refuses_a_batch_above_its_scratch(gpu_ctx):
    count = NOALLOC_MAX // 40 + 1000       # 24 + 16 descriptor bytes per 1x1 matrix
    buf = gpu_ctx.alloc(count)
    ptrs = [buf.offset(j) for j in range(count)]
    rc, _keep = _call("fap_fw_batched_noalloc", gpu_ctx, ptrs, np.ones(count), np.ones(count))
    assert rc == -1                        # FAP_E_ARG
    msg = fa.load().fap_last_error().decode()
    assert "scratch" in msg and str(NOALLOC_MAX) in msg
    # nothing was launched: the library still works on the same device
    m = _mats([70], 9)
    a = _upload(gpu_ctx, m)
    rc, _keep = _call("fap_fw_batched_noalloc", gpu_ctx, a[1], [70], a[2])
    assert rc == 0
    np.testing.assert_array_equal(_read(a[0], m, a[2], a[3])[0], O.floyd(m[0]))
    a[0].free()
    buf.free()
