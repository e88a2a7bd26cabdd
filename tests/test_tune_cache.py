"""The autotuner's cache on the host (include/fap_tune.h): buckets, value
checks, the text file round trip and its error handling.  No device."""

import os

import numpy as np
import pytest

import fastapsp_amd as fa
from fastapsp_amd import TuneCache


def test_bucket_keys():
    kf = TuneCache.key_fw
    # quarter octaves: smallest k with n^4 <= 2^k
    assert [kf(n) for n in (1, 2, 3, 64, 65, 4096, 5120, 6144, 8192, 8193)] == \
        [0, 4, 7, 24, 25, 48, 50, 51, 52, 53]
    # (V bucket * 64 + round(2 log2(1 + E/V))) * 32 + sources bucket
    db = round(2 * np.log2(1 + 952362 / 262144))
    assert TuneCache.key_sssp(262144, 952362, 74) == (18 * 64 + db) * 32 + 7
    assert TuneCache.key_sssp(262144, 952362, 3669) == (18 * 64 + db) * 32 + 12
    assert TuneCache.key_sssp(4096, 0, 1) == 12 * 64 * 32
    assert TuneCache.key_sssp(1, 0, 1) == 0
    # 64-tile triples: one 65-vertex job has 2^3 = 8, plus a 64-vertex job's 1
    assert TuneCache.key_batch([65, 64]) == 4
    assert TuneCache.key_batch([]) == 0


def test_set_get_and_value_checks():
    c = TuneCache()
    assert c.get("fw", 13) is None and len(c) == 0
    c.set("fw", 13, [128, 1], 21.5)
    c.set("sssp", 1154, [512, 16], 9.25)
    c.set("batch", 21, [8], 25.0)
    assert c.get("fw", 13) == ((128, 1), 21.5)
    assert c.get("sssp", 1154) == ((512, 16), 9.25)
    assert c.get("batch", 21) == ((8,), 25.0)
    c.set("fw", 13, [64, 0], 20.0)              # replaces
    assert c.get("fw", 13)[0] == (64, 0) and len(c) == 3
    for kind, p in (("fw", [96, 0]), ("fw", [64, 2]), ("sssp", [300, 8]), ("sssp", [256, 0]),
                    ("batch", [0]), ("batch", [65]), ("fw", [64]), ("nope", [1])):
        with pytest.raises(fa.FapError):
            c.set(kind, 1, p, 1.0)
    with pytest.raises(fa.FapError):
        c.set("fw", -1, [64, 0], 1.0)


def test_file_round_trip_and_merge(tmp_path):
    c = TuneCache()
    c.set("fw", 13, [128, 0], 22.97)
    c.set("sssp", 1154, [1024, 8], 14.5625)
    c.set("batch", 21, [4], 25.277)
    p = str(tmp_path / "tune.txt")
    c.save(p)
    d = TuneCache(p)
    assert len(d) == 3
    assert d.get("fw", 13) == ((128, 0), pytest.approx(22.97))
    assert d.get("batch", 21)[0] == (4,)
    # a second file merges, later entries win
    q = str(tmp_path / "more.txt")
    with open(q, "w") as f:
        f.write("# comment\n\nfw 13 64 1 20.0  # trailing comment\nfw 10 64 0 1.5\n")
    d.load(q)
    assert len(d) == 4 and d.get("fw", 13)[0] == (64, 1)


@pytest.mark.parametrize("line", ["fw 13 96 0 1.0", "fw 13 64 0", "fw x 64 0 1.0",
                                  "sssp 5 256 8 1.0 extra", "batch 3 0 1.0", "zz 1 1 1.0",
                                  "fw 13 64 0 -1"])
def test_malformed_file_is_rejected_whole(tmp_path, line):
    p = str(tmp_path / "bad.txt")
    with open(p, "w") as f:
        f.write("fw 12 64 0 3.0\n" + line + "\n")
    c = TuneCache()
    with pytest.raises(fa.FapError) as e:
        c.load(p)
    assert e.value.code == -5 and ":2:" in str(e.value)
    assert len(c) == 0                          # nothing merged from a bad file


def test_missing_file(tmp_path):
    with pytest.raises(fa.FapError) as e:
        TuneCache(str(tmp_path / "absent.txt"))
    assert e.value.code == -5
    assert not os.path.exists(str(tmp_path / "absent.txt"))
