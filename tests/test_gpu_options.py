"""Context launch options (fap_ctx_set_option / fap_ctx_get_option, listed
in include/fap_dev.h): every documented option takes its valid values and
reads them back, and refuses values outside its range and unknown names
with FAP_E_ARG, leaving the value it had; a fresh context starts at the
documented defaults.  Runs on its own context, not the session's."""

import pytest

import fastapsp_amd as fa
from fastapsp_amd import FapError

pytestmark = pytest.mark.gpu

FAP_E_ARG = -1

# name: (default as read back, valid values, invalid values)
OPTIONS = {
    "fw_lookahead": (-1, [-1, 0, 1], [-2, 2]),
    "fw_tile": (0, [0, 64, 128], [32, 256, -1]),
    "sssp_threads": (0, [0, 256, 512, 1024], [128, 2048]),
    "sssp_delta": (0, [0, 1, 5000], [-1]),
    "sssp_group": (0, [0, 1, 4, 8], [2, 16]),
    "sssp_bin": (3, [0, 1, 2, 3], [-1, 4]),
    "sssp_wgs": (0, [0, 1, 8], [-1, 9]),
    "sssp_index": (1, [0, 1], [2, -1]),
    "bidx_tile": (0, [0, 64, 128], [32, 256]),
    "bidx_delta": (0, [0, 7, 100000], [-1]),
    "batch_grid": (0, [0, 1, 64], [-1, 65]),
    "batch_wait_us": (0, [0, 1, 2000000], [-1]),
}


def test_options_round_trip_and_reject_bad_values(gpu_ctx):
    ctx = fa.Context(0)
    try:
        for name, (default, good, bad) in OPTIONS.items():
            assert ctx.get_option(name) == default, name
            for v in good:
                ctx.set_option(name, v)
                assert ctx.get_option(name) == v, (name, v)
            for v in bad:
                with pytest.raises(FapError) as e:
                    ctx.set_option(name, v)
                assert e.value.code == FAP_E_ARG, (name, v)
                assert ctx.get_option(name) == good[-1], f"{name} changed by a refused value"
        # auto values that resolve on set
        assert ctx.get_option("sssp_pack") == 1 and ctx.get_option("sssp_arcs") == 2
        ctx.set_option("sssp_pack", 2)
        ctx.set_option("sssp_pack", 0)
        assert ctx.get_option("sssp_pack") == 1
        ctx.set_option("sssp_arcs", 4)
        ctx.set_option("sssp_arcs", 0)
        assert ctx.get_option("sssp_arcs") == 2
        with pytest.raises(FapError) as e:
            ctx.set_option("no_such_option", 1)
        assert e.value.code == FAP_E_ARG
        with pytest.raises(FapError):
            ctx.get_option("no_such_option")
    finally:
        ctx.close()
