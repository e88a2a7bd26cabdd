"""The drop-in claim for C++ callers: a program written against the
reference's fap/fap.h API compiles unchanged against include/fap/fap.h and
links libfap_amd.so (CPU); run on the GPU it solves .mtx graphs the way the
reference's example does (app/singleNodeExample.cpp:40-68) and its own
check_result passes."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from fastapsp_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "fastapsp_amd", "_lib")
SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_example.cpp")


def build(tmp_path):
    exe = str(tmp_path / "drop_in_example")
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    cmd = [cxx, "-std=c++17", "-O1", SRC, "-I" + os.path.join(ROOT, "include"),
           "-L" + LIBDIR, "-lfap_amd", "-Wl,-rpath," + LIBDIR, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_example_compiles_and_links_against_the_drop_in(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def _write_grid(tmp_path, R, drop, diag, wl, wh, seed):
    u, v, w = synth.grid_graph(R, R, drop, diag, wl, wh, seed)
    name = f"g{R}"
    synth.write_mtx(str(tmp_path / "graph" / "weight-undirected" / f"{name}.mtx"), R * R, u, v, w)
    return name


@pytest.mark.gpu
@pytest.mark.parametrize("R,K", [(64, 8), (150, 24)])
def test_example_runs_and_verifies(tmp_path, R, K):
    exe = build(tmp_path)
    name = _write_grid(tmp_path, R, 0.05, 0.02, 1, 50, R)
    r = subprocess.run([exe, name, str(K), "0", "1"], capture_output=True, text=True,
                       cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    want_id = 3 if R * R > 20000 else K
    assert f"part {want_id} ok" in r.stdout
    if R * R > 20000:
        assert "(0 path walks)" not in r.stdout
