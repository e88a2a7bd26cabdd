"""The autotuner on the GPU: each tuner records a valid entry for the
workload's bucket, a context follows its cache for options left at auto,
and results are the oracle's under every setting the tuners can pick."""

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import TuneCache, synth

pytestmark = pytest.mark.gpu


def dense(n, seed):
    return synth.dense_matrix(n, 0.02, 1, 1000, seed)


def test_tuners_record_entries(tmp_path):
    with fa.Context(0) as ctx:
        with pytest.raises(fa.FapError):
            ctx.tune_fw(200)                      # no cache attached
        cache = TuneCache()
        ctx.set_tune_cache(cache)
        ctx.tune_fw(300, reps=1)
        p, ms = cache.get("fw", TuneCache.key_fw(300))
        assert p[0] in (64, 128) and p[1] in (0, 1) and ms > 0
        u, v, w = synth.grid_graph(40, 40, 0.1, 0.02, 1, 100, 2)
        ro, col, wt = synth.to_csr(1600, u, v, w, False)
        dg = ctx.upload_graph(ro, col, wt)
        ctx.tune_sssp(dg, np.arange(0, 1600, 50), reps=1)
        p, ms = cache.get("sssp", TuneCache.key_sssp(1600, len(col), 32))
        assert p[0] in (256, 512, 1024) and p[1] in (2, 4, 8, 16, 32)
        sizes = [70, 130, 200, 64, 1]
        ctx.tune_batch(sizes, reps=1)
        p, _ = cache.get("batch", TuneCache.key_batch(sizes))
        assert p[0] in (4, 8, 16, 32, 64)
        # the options were restored: still auto
        assert ctx.get_option("fw_tile") == 0 and ctx.get_option("fw_lookahead") == -1
        assert ctx.get_option("sssp_threads") == 0 and ctx.get_option("batch_grid") == 0
        cache.save(str(tmp_path / "t.txt"))
        assert len(TuneCache(str(tmp_path / "t.txt"))) == 3
        dg.close()


@pytest.mark.parametrize("params", [(64, 0), (64, 1), (128, 0), (128, 1)])
def test_fw_follows_cache_and_stays_exact(params):
    n = 400
    A = dense(n, 21)
    want = O.floyd(A)
    with fa.Context(0) as ctx:
        cache = TuneCache()
        cache.set("fw", TuneCache.key_fw(n), list(params), 1.0)
        ctx.set_tune_cache(cache)
        d = ctx.to_device(A)
        ctx.fw(d.ptr, n, n)
        np.testing.assert_array_equal(d.get(), want)


@pytest.mark.parametrize("params", [(256, 2), (512, 32), (1024, 4)])
def test_sssp_follows_cache_and_stays_exact(params):
    u, v, w = synth.grid_graph(30, 30, 0.1, 0.02, 1, 100, 8)
    ro, col, wt = synth.to_csr(900, u, v, w, False)
    src = np.array([0, 17, 450, 899], np.int32)
    with fa.Context(0) as ctx:
        cache = TuneCache()
        cache.set("sssp", TuneCache.key_sssp(900, len(col), 4), list(params), 1.0)
        ctx.set_tune_cache(cache)
        dg = ctx.upload_graph(ro, col, wt)
        ds = ctx.to_device(src)
        out = ctx.alloc((4, 900))
        dg.sssp(ds.ptr, 4, out.ptr, 900)
        np.testing.assert_array_equal(out.get(), O.batched_sssp(src, ro, col, wt))
        dg.close()


@pytest.mark.parametrize("per_cu", [1, 16, 64])
def test_batch_follows_cache_and_stays_exact(per_cu):
    sizes = [70, 130, 200, 64, 1, 129]
    mats = [dense(n, 40 + i) for i, n in enumerate(sizes)]
    with fa.Context(0) as ctx:
        cache = TuneCache()
        cache.set("batch", TuneCache.key_batch(sizes), [per_cu], 1.0)
        ctx.set_tune_cache(cache)
        ds = [ctx.to_device(m) for m in mats]
        ctx.fw_batched([d.ptr for d in ds], sizes, sizes)
        for d, m in zip(ds, mats):
            np.testing.assert_array_equal(d.get(), O.floyd(m))


def test_probe_minplus_ceiling(gpu_ctx):
    """The instruction-mix probe runs and reports a sane rate: on an idle
    MI355X ~76 TOP/s (v_add_u32 issues at twice the rate of v_min3_i32);
    the bounds leave room for clock changes and other work on the device."""
    tops = gpu_ctx.probe_minplus_ceiling()
    assert 20.0 < tops < 100.0, tops
