"""Host side of the drop-in, on the CPU (no GPU calls): the C ABI library
loads and exports every declared symbol, the .mtx reader reproduces the
reference's CSR order, preCondition reproduces findBoundry/sort_and_encode,
the built-in partitioner is balanced and deterministic, and bad input is
reported as an error code."""

import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import fastapsp_amd as fa
from fastapsp_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_GRAPH_DIR = "/root/reference/app"


def declared_symbols():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fap\w*)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = fa.load()
    syms = declared_symbols()
    assert len(syms) > 50
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python bindings cover them
    from fastapsp_amd import _SIGS
    assert not (syms - set(_SIGS)), sorted(syms - set(_SIGS))


def test_no_gpu_calls_report_errors_cleanly():
    lib = fa.load()
    assert lib.fap_version().startswith(b"fap-mi355x")
    # null arguments are rejected before any HIP call
    assert lib.fap_minplus(None, None, 1, None, 1, None, 1, 1, 1, 1, 0, None) == -1
    assert "null" in fa.last_error()


def csr_of(R, drop, diag, wl, wh, seed):
    u, v, w = synth.grid_graph(R, R, drop, diag, wl, wh, seed)
    V = R * R
    return V, u, v, w, synth.to_csr(V, u, v, w, False)


@pytest.mark.parametrize("weighted,directed", [(False, False), (True, False), (True, True),
                                               (False, True)])
def test_reader_matches_reference_order(tmp_path, monkeypatch, weighted, directed):
    V, u, v, w, _ = csr_of(12, 0.1, 0.05, 1, 40, 3)
    sub = ("weight-" if weighted else "unweight-") + ("directed" if directed else "undirected")
    synth.write_mtx(str(tmp_path / "graph" / sub / "t.mtx"), V, u, v, w if weighted else None)
    # a leading comment line, as the reader skips them
    p = tmp_path / "graph" / sub / "t.mtx"
    p.write_text("%%MatrixMarket test\n% comment\n" + p.read_text())
    monkeypatch.chdir(tmp_path)
    with fa.Graph.from_file("t", directed, weighted, 3) as g:
        ro, col, wt = g.csr()
    ww = w if weighted else np.ones_like(w)
    e_ro, e_col, e_w = synth.csr_reference_order(V, u, v, ww, directed)
    np.testing.assert_array_equal(ro, e_ro)
    np.testing.assert_array_equal(col, e_col)
    np.testing.assert_array_equal(wt, e_w)
    if O.ref_available():
        _, r_ro, r_col, r_w = O.ref_read("t", directed, weighted)
        np.testing.assert_array_equal(ro, r_ro)
        np.testing.assert_array_equal(col, r_col[: len(col)])
        np.testing.assert_array_equal(wt, r_w[: len(col)].astype(np.int32))


@pytest.mark.skipif(not (O.ref_available() and os.path.isdir(REF_GRAPH_DIR + "/graph")),
                    reason="needs the reference's fixture graph and oracle/_ref")
def test_reader_on_reference_fixture(monkeypatch):
    monkeypatch.chdir(REF_GRAPH_DIR)
    with fa.Graph.from_file("luxembourg_osm", False, False, 128) as g:
        ro, col, wt = g.csr()
        assert g.num_vertexs == 114599
    _, r_ro, r_col, r_w = O.ref_read("luxembourg_osm", False, False)
    np.testing.assert_array_equal(ro, r_ro)
    np.testing.assert_array_equal(col, r_col)
    assert (wt == 1).all() and (r_w == 1).all()


@pytest.mark.parametrize("K", [1, 2, 7, 8])
def test_precondition_matches_oracle_for_given_partition(K):
    V, u, v, w, (ro, col, wt) = csr_of(18, 0.1, 0.03, 1, 9, K)
    gid = np.random.default_rng(K).integers(1, K + 1, V).astype(np.int32)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        g.setPartition(gid)
        g.preCondition()
        pc = O.precondition(K, gid, ro, col)
        np.testing.assert_array_equal(g.getGraphId(), gid)
        np.testing.assert_array_equal(g.getMapping(), pc["st2ed"])
        np.testing.assert_array_equal(g.boundary_flags(), pc["is_bdy"])
        vn, vo, bn, bo = g.part_info()
        for a, b in ((vn, pc["ver_num"]), (vo, pc["ver_off"]), (bn, pc["bdy_num"]),
                     (bo, pc["bdy_off"])):
            np.testing.assert_array_equal(a[1:], b[1:])
        assert g.isSplit() == (V > 20000)


def test_partitioner_balanced_deterministic_and_compact():
    V, u, v, w, (ro, col, wt) = csr_of(64, 0.05, 0.0, 1, 1, 1)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), 8) as g1, \
            fa.Graph.from_csr(ro, col, wt.astype(np.float32), 8) as g2:
        a, b = g1.getGraphId(), g2.getGraphId()
        np.testing.assert_array_equal(a, b)
        assert g1.used_metis() is False
        counts = np.bincount(a, minlength=9)[1:]
        assert counts.min() >= 500 and counts.max() <= 524
        g1.preCondition()
        _, _, bn, _ = g1.part_info()
        # a 64x64 grid cut 8 ways: a compact cut keeps boundaries near the perimeter
        assert bn[1:].max() < 110


@pytest.mark.parametrize("K", [3, 5, 16])
def test_partitioner_handles_odd_K_and_components(K):
    # two disconnected grids plus isolated vertices
    V1, u1, v1, w1, _ = csr_of(10, 0.0, 0.0, 1, 1, 1)
    u = np.concatenate([u1, u1 + V1])
    v = np.concatenate([v1, v1 + V1])
    V = 2 * V1 + 5
    ro, col, wt = synth.to_csr(V, u, v, np.ones_like(u), False)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        gid = g.getGraphId()
        assert gid.min() >= 1 and gid.max() <= K
        counts = np.bincount(gid, minlength=K + 1)[1:]
        assert counts.sum() == V and counts.max() - counts.min() <= 2


def test_bad_inputs_are_errors(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with pytest.raises(fa.FapError) as e:
        fa.Graph.from_file("missing", False, False, 2)
    assert e.value.code == -5
    d = tmp_path / "graph" / "weight-undirected"
    d.mkdir(parents=True)
    (d / "bad.mtx").write_text("3 3 2\n1 2 1.5\n2 3 2\n")
    with pytest.raises(fa.FapError):
        fa.Graph.from_file("bad", False, True, 2)   # non-integral weight
    (d / "oob.mtx").write_text("3 3 2\n1 2 1\n2 9 2\n")
    with pytest.raises(fa.FapError):
        fa.Graph.from_file("oob", False, True, 2)   # vertex id out of range
    (d / "ok.mtx").write_text("3 3 2\n1 2 1\n2 3 2\n")
    with pytest.raises(fa.FapError):
        fa.Graph.from_file("ok", False, True, 4)    # K > V
    with pytest.raises(fa.FapError):
        fa.Graph.from_csr(np.array([0, 1, 2]), np.array([1, 5]), np.array([1.0, 1.0]), 1)
    g = fa.Graph.from_file("ok", False, True, 2)
    with pytest.raises(fa.FapError):
        g.setPartition(np.array([1, 2, 3], np.int32))  # id > K
    with pytest.raises(fa.FapError):
        g.solve()                                      # before preCondition
    g.close()


def test_bench_workloads_are_defined():
    """Every --workload choice of bench.py dispatches to a defined function,
    and --gpus above 1 is refused with a message (no device touched)."""
    import importlib.util
    import subprocess
    import sys
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for name in ("dense", "copy", "c1", "grid_pipeline", "grid_batch", "fw_sweep", "sssp_sweep",
                 "batch", "serve"):
        assert callable(getattr(bench, "bench_" + name)), name
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_batched_closure_schedule_simulated_on_cpu(tmp_path):
    """The batched closure's plan builder (from the library) and the task ->
    tile maps its kernels use, run as a CPU simulation over ragged batches
    (tests/cpp/batch_schedule_check.cpp): tiles written once per launch, no
    launch reads a tile it writes elsewhere, one kernel instance per launch,
    and every matrix equal to plain Floyd-Warshall.  The same program models
    one persistent launch with in-kernel waits on per-job phase counters
    under in-order dispatch into R resident slots: a static task deal with
    more workgroups than slots must be found stuck, a bounded wait must end
    it with the error flag, and in-order task claims (also claimed one task
    ahead, as fw_queue_kernel does) over the library's own single-launch
    queue (fapi::build_fw_queue, checked against the launch sequence) must
    drain for any grid, with every matrix equal to plain Floyd-Warshall."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("no hipcc")
    libdir = os.path.join(ROOT, "fastapsp_amd", "_lib")
    exe = str(tmp_path / "batch_schedule_check")
    cmd = [hipcc, "-std=c++17", "-O2", os.path.join(ROOT, "tests", "cpp", "batch_schedule_check.cpp"),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "fastapsp_amd", "csrc"),
           "-L" + libdir, "-lfap_amd", "-Wl,-rpath," + libdir, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


def _build_metis_tools(tmp_path):
    import shutil
    import subprocess
    gcc, gxx = shutil.which("gcc"), shutil.which("g++")
    if not gcc or not gxx:
        pytest.skip("no gcc/g++")
    lib = str(tmp_path / "libmetis_recorder.so")
    ref = str(tmp_path / "metis_input_ref")
    r = subprocess.run([gcc, "-shared", "-fPIC", "-O1", "-o", lib,
                        os.path.join(ROOT, "tests", "cpp", "metis_recorder.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-o", ref,
                        os.path.join(ROOT, "tests", "cpp", "metis_input_ref.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib, ref


def _reference_metis_input(ref, directed, ro, col):
    import subprocess
    text = f"{int(directed)} {len(ro) - 1} {len(col)}\n" + " ".join(map(str, ro)) + "\n" + \
        " ".join(map(str, col)) + "\n"
    r = subprocess.run([ref], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    return [int(x) for x in lines[0].split()], [int(x) for x in lines[1].split()]


@pytest.mark.parametrize("directed", [False, True])
def test_metis_adapter_hands_metis_what_graph_decomposition_builds(tmp_path, monkeypatch,
                                                                   directed):
    """The METIS adapter (partition.cpp) against a recording stand-in for
    libmetis (tests/cpp/metis_recorder.c): METIS_PartGraphKway receives
    exactly the xadj / adjncy the reference's Graph_decomposition builds
    (fap/utils/readIdFile.h:71-112, restated in tests/cpp/metis_input_ref.cpp
    with the same std::map / std::unordered_set, so neighbour order matches
    too), ncon 1, nparts K and NULL for every optional argument
    (readIdFile.h:118-121); the returned part[i] becomes graph id
    part[i] + 1 (readIdFile.h:127-131)."""
    lib, ref = _build_metis_tools(tmp_path)
    rec = str(tmp_path / "record.txt")
    monkeypatch.setenv("FAP_METIS_LIB", lib)
    monkeypatch.setenv("FAP_METIS_RECORD", rec)
    monkeypatch.delenv("FAP_NO_METIS", raising=False)
    rng = np.random.default_rng(3)
    V, K = 300, 7
    if directed:
        # every vertex on a directed cycle (no isolated vertex), random chords, a self loop
        u = np.r_[np.arange(V), rng.integers(0, V, 700), 5]
        v = np.r_[(np.arange(V) + 1) % V, rng.integers(0, V, 700), 5]
    else:
        a, b = rng.integers(0, V, 800), rng.integers(0, V, 800)
        keep = a != b
        u, v = a[keep], b[keep]     # both arcs of each edge added by to_csr
    w = rng.integers(1, 50, len(u))
    ro, col, wt = synth.to_csr(V, u, v, w, directed)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        assert g.used_metis() is True
        assert g.directed is directed
        gid = g.getGraphId()
    lines = open(rec).read().split("\n")
    nv, ncon, nparts = map(int, lines[0].split())
    assert (nv, ncon, nparts) == (V, 1, K)
    assert [int(x) for x in lines[1].split()] == [1, 1, 1, 1, 1, 1]   # all NULL
    xadj = [int(x) for x in lines[2].split()]
    adjncy = [int(x) for x in lines[3].split()]
    want_x, want_a = _reference_metis_input(ref, directed, ro, col)
    assert xadj == want_x
    assert adjncy == want_a
    np.testing.assert_array_equal(gid, (7 * np.arange(V)) % K + 1)


def test_metis_adapter_refuses_directed_graph_with_isolated_vertex(tmp_path, monkeypatch):
    """Documented difference: for a directed graph with a vertex that has no
    arc (other than a self loop), the reference's symmetrised xadj has fewer
    than V + 1 entries while it still passes nVertices = V
    (readIdFile.h:80-112), so METIS reads past the array; the adapter
    refuses such a graph with an error instead, and METIS is not called."""
    lib, _ = _build_metis_tools(tmp_path)
    rec = str(tmp_path / "record.txt")
    monkeypatch.setenv("FAP_METIS_LIB", lib)
    monkeypatch.setenv("FAP_METIS_RECORD", rec)
    monkeypatch.delenv("FAP_NO_METIS", raising=False)
    V = 10
    u = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8])     # vertex 9 has no arc
    v = np.array([1, 2, 3, 4, 5, 6, 7, 8, 0])
    ro, col, wt = synth.to_csr(V, u, v, np.ones_like(u), True)
    with pytest.raises(fa.FapError, match="without arcs"):
        g = fa.Graph.from_csr(ro, col, wt.astype(np.float32), 3)
        g.close()
    assert not os.path.exists(rec)


def test_host_code_under_address_and_undefined_sanitizers(tmp_path):
    """The library's host sources (reader, partitioners, preprocessing,
    fapGraph and its C ABI, tune cache) built with
    -fsanitize=address,undefined and driven by tests/cpp/host_sanitize.cpp
    over valid and malformed inputs (and the METIS adapter through the
    stand-in libmetis); the kernels' objects are linked as built, no GPU
    call is made.  Any sanitizer report fails the run."""
    import glob
    import shutil
    import subprocess
    import concurrent.futures as cf
    gxx, gcc = shutil.which("g++"), shutil.which("gcc")
    hip_objs = sorted(glob.glob(os.path.join(ROOT, "fastapsp_amd", "_build", "*.hip.o")))
    if not gxx or not gcc or not hip_objs:
        pytest.skip("needs g++/gcc and the built kernel objects (fastapsp_amd/build.py)")
    csrc = os.path.join(ROOT, "fastapsp_amd", "csrc")
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
             "-I" + csrc]
    srcs = [os.path.join(csrc, f) for f in sorted(os.listdir(csrc)) if f.endswith(".cpp")]
    srcs.append(os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"))

    def compile_one(src):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        r = subprocess.run([gxx] + flags + ["-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with cf.ThreadPoolExecutor(4) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = str(tmp_path / "host_sanitize")
    r = subprocess.run([gxx, "-fsanitize=address,undefined", "-o", exe] + objs + hip_objs +
                       ["-L/opt/rocm/lib", "-lamdhip64", "-ldl", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    rec = str(tmp_path / "libmetis_recorder.so")
    r = subprocess.run([gcc, "-shared", "-fPIC", "-o", rec,
                        os.path.join(ROOT, "tests", "cpp", "metis_recorder.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("FAP_NO_METIS", None)
    r = subprocess.run([exe, str(scratch), rec], capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout + r.stderr)[-4000:]
    assert "runtime error" not in r.stderr, r.stderr[-4000:]


def test_boundary_index_host_code_with_emulated_kernels_under_sanitizers(tmp_path):
    """The boundary index's host code (fastapsp_amd/csrc/bindex.hip: index
    build, border order and lists, the block iterations' workspace and
    bookkeeping), compiled as it is against a host-only stand-in of the HIP
    runtime and CPU emulations of its kernels (tests/cpp/bidx_host), under
    -fsanitize=address,undefined: a 64 x 64 grid cut in 16 parts, boundary
    rows for calls of 4 and of all 16 parts, sampled rows equal to Dijkstra."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    R, K = 64, 16
    V = R * R
    u, v, w = synth.grid_graph(R, R, 0.10, 0.02, 1, 100, 3)
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    with fa.Graph.from_csr(ro, col, wt.astype(np.float32), K) as g:
        g.preCondition()
        vn, vo, bn, _ = g.part_info()
        st2ed, gid = g.getMapping(), g.getGraphId()
    ed2st = np.empty(V, np.int32)
    ed2st[st2ed] = np.arange(V, dtype=np.int32)
    gpath = str(tmp_path / "grid.bin")
    with open(gpath, "wb") as f:
        np.array([V, len(col), K], np.int64).tofile(f)
        for a in (ro, col, wt, gid, st2ed, ed2st, vn, vo, bn):
            np.asarray(a, np.int32).tofile(f)
    src = open(os.path.join(ROOT, "fastapsp_amd", "csrc", "bindex.hip")).read()
    a = src.index("struct BidxPart {")
    b = src.index("namespace fapk {")
    c = src.index("namespace fapi {\n\nnamespace {\n\nvoid free_bidx")
    (tmp_path / "bindex_host.inc").write_text(src[a:b] + '\n#include "emu_kernels.h"\n' + src[c:])
    here = os.path.join(ROOT, "tests", "cpp", "bidx_host")
    inc = ["-I" + os.path.join(here, "mock"), "-I" + here, "-I" + str(tmp_path),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "fastapsp_amd", "csrc")]
    fw_o = str(tmp_path / "fw_emu.o")
    r = subprocess.run([gxx, "-std=c++17", "-O2", "-c", os.path.join(here, "fw_emu.cpp"), "-o", fw_o],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = str(tmp_path / "bidx_host")
    r = subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=undefined"] + inc +
                       [os.path.join(here, "harness.cpp"), fw_o, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    for args in (["4"], ["16", "6"]):     # calls of 4 parts; one call of 16 (6 row buffers)
        r = subprocess.run([exe, gpath] + args, capture_output=True, text=True, timeout=600,
                           env=env)
        assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
        assert "0 bad" in r.stdout and "runtime error" not in r.stderr, r.stdout + r.stderr
