"""Pins the CPU oracle (oracle/fap_oracle.cpp) before anything is checked
against it.  Every oracle function is checked by something other than its own
formula:

* golden vectors from scipy's Dijkstra / numpy brute force (tests/golden/),
* closed forms (unit grid = Manhattan distance, weighted path = w*|i-j|,
  disconnected = INF),
* an independent algorithm on random graphs (vectorised Bellman-Ford),
* the reference's own CPU code compiled from its headers (oracle/_ref), where
  that library has been built,
* partition invariance: a part block equals the Dijkstra rows of its vertices.
"""

import os

import numpy as np
import pytest

import oracle_lib as O
from fastapsp_amd import synth

G = os.path.join(os.path.dirname(__file__), "golden")
INF = O.INF


def golden(name):
    return np.load(os.path.join(G, name))


def bellman_ford_all(V, ro, col, w):
    """Independent APSP: synchronous Bellman-Ford on all sources at once."""
    src = np.repeat(np.arange(V), np.diff(ro))
    d = np.full((V, V), INF, np.int64)
    d[np.arange(V), np.arange(V)] = 0
    for _ in range(V):
        cand = d[:, src] + w[None, :]
        new = d.copy()
        np.minimum.at(new.T, col, cand.T)
        new = np.minimum(new, INF)
        if np.array_equal(new, d):
            break
        d = new
    return d.astype(np.int32)


def dense_of(V, ro, col, w):
    m = np.full((V, V), INF, np.int32)
    src = np.repeat(np.arange(V), np.diff(ro))
    m[src, col] = w
    m[np.arange(V), np.arange(V)] = 0
    return m


# ---------- golden vectors ----------

@pytest.mark.parametrize("name", ["grid8.npz", "rgg100.npz"])
def test_dijkstra_matches_scipy_golden(name):
    g = golden(name)
    ro, col, w, apsp = g["row_offset"], g["col"], g["wt"], g["apsp"]
    V = len(ro) - 1
    got = O.batched_sssp(np.arange(V), ro, col, w)
    np.testing.assert_array_equal(got, apsp)


@pytest.mark.parametrize("name", ["grid8.npz", "rgg100.npz"])
def test_floyd_of_adjacency_matches_scipy_golden(name):
    g = golden(name)
    ro, col, w, apsp = g["row_offset"], g["col"], g["wt"], g["apsp"]
    np.testing.assert_array_equal(O.floyd(dense_of(len(ro) - 1, ro, col, w)), apsp)


def test_floyd_dense_matches_scipy_golden():
    g = golden("dense70.npz")
    np.testing.assert_array_equal(O.floyd(g["mat"]), g["closure"])


def test_min_plus_matches_bruteforce_golden():
    g = golden("minplus_13x37x29.npz")
    np.testing.assert_array_equal(O.min_plus(g["A"], g["B"]), g["C"])


def test_golden_check_catches_planted_mistakes():
    """The golden comparison is sensitive to the mistakes it must catch."""
    g = golden("minplus_13x37x29.npz")
    A, B, Cg = g["A"], g["B"], g["C"]
    # a dropped term (last k ignored)
    assert not np.array_equal(O.min_plus(A[:, :-1], B[:-1, :]), Cg)
    # a transposed operand (same shapes only for square pieces)
    As = A[:, :13]
    assert not np.array_equal(O.min_plus(As, B[:13, :]), O.min_plus(As.T.copy(), B[:13, :]))
    # a wrong weight in Floyd-Warshall input
    d = golden("dense70.npz")
    m = d["mat"].copy()
    i, j = np.argwhere((m > 0) & (m < INF))[0]
    m[i, j] += 1
    assert not np.array_equal(O.floyd(m), d["closure"]) or d["closure"][i, j] < d["mat"][i, j]


@pytest.mark.skipif(not os.path.exists(os.path.join(G, "luxembourg_rows.npz"))
                    or not os.path.exists("/root/reference/app/graph"),
                    reason="needs the reference's fixture graph")
def test_dijkstra_on_reference_fixture_graph():
    d = golden("luxembourg_rows.npz")
    path = "/root/reference/app/graph/unweight-undirected/luxembourg_osm.mtx"
    with open(path) as f:
        V = int(f.readline().split()[0])
        e = np.loadtxt(f, dtype=np.int64).reshape(-1, 2) - 1
    ro, col, w = synth.to_csr(V, e[:, 0], e[:, 1], np.ones(len(e), np.int64), False)
    got = O.batched_sssp(d["sources"], ro, col, w)
    np.testing.assert_array_equal(got, d["rows"])


def test_fixture_graph_copy_gives_the_golden_rows():
    """tests/golden/luxembourg_osm.npz (the reference's fixture graph kept as
    its edge lines) reproduces the two scipy Dijkstra rows taken from the
    original file, so the copy can stand in for it where the reference
    checkout is absent."""
    z, d = golden("luxembourg_osm.npz"), golden("luxembourg_rows.npz")
    V, e = int(z["V"]), z["edges"].astype(np.int64) - 1
    assert int(z["lines"]) == len(e) == 119666 and V == int(d["V"]) == 114599
    ro, col, w = synth.to_csr(V, e[:, 0], e[:, 1], np.ones(len(e), np.int64), False)
    np.testing.assert_array_equal(O.batched_sssp(d["sources"], ro, col, w), d["rows"])


# ---------- closed forms ----------

def test_unit_grid_is_manhattan():
    R = 10
    u, v, w = synth.grid_graph(R, R, 0.0, 0.0, 1, 1, 0)
    ro, col, wt = synth.to_csr(R * R, u, v, w, False)
    d = O.batched_sssp(np.arange(R * R), ro, col, wt)
    r, c = np.divmod(np.arange(R * R), R)
    manh = np.abs(r[:, None] - r[None, :]) + np.abs(c[:, None] - c[None, :])
    np.testing.assert_array_equal(d, manh)
    np.testing.assert_array_equal(O.floyd(dense_of(R * R, ro, col, wt)), manh)


def test_weighted_path_and_components():
    n, wgt = 9, 7
    u = np.arange(n - 1)
    v = u + 1
    w = np.full(n - 1, wgt)
    # second component: 3 isolated vertices appended
    V = n + 3
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    d = O.batched_sssp(np.arange(V), ro, col, wt)
    i = np.arange(n)
    np.testing.assert_array_equal(d[:n, :n], wgt * np.abs(i[:, None] - i[None, :]))
    assert (d[:n, n:] == INF).all() and (d[n:, :n] == INF).all()
    assert (np.diag(d) == 0).all()


# ---------- independent algorithm ----------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dijkstra_and_floyd_vs_bellman_ford(seed):
    n = 40
    au, av, aw = synth.rgg_graph(n, 4, 0, 20, seed)  # weight 0 allowed
    ro, col, w = synth.to_csr(n, au, av, aw, True)
    bf = bellman_ford_all(n, ro.astype(np.int64), col, w.astype(np.int64))
    np.testing.assert_array_equal(O.batched_sssp(np.arange(n), ro, col, w), bf)
    np.testing.assert_array_equal(O.floyd(dense_of(n, ro, col, w)), bf)


def test_min_plus_path_vs_numpy_first_argmin():
    """min_plus_path_CPU_advanced's label is B_path[k][j] of the FIRST k
    attaining the minimum (strict relaxation in k order,
    fap/kernel/minplus.h:51-65), -1 where the result is INF.  Small weights
    force ties, so a last-k or any-k rule fails."""
    rng = np.random.default_rng(3)
    for M, N, K in ((17, 23, 9), (40, 31, 64), (5, 7, 0)):
        A = rng.integers(0, 4, (M, K)).astype(np.int32)
        B = rng.integers(0, 4, (K, N)).astype(np.int32)
        A[rng.random((M, K)) < 0.3] = O.INF
        B[rng.random((K, N)) < 0.3] = O.INF
        A[0, :] = O.INF                       # an all-INF row
        Bp = rng.integers(-5, 1000, (K, N)).astype(np.int32)
        Cm, Cp = O.min_plus(A, B, Bp)
        if K:
            S = A[:, :, None].astype(np.int64) + B[None, :, :]
            want = np.minimum(S.min(1), O.INF)
            k = S.argmin(1)
            wantp = np.where(want < O.INF, Bp[k, np.arange(N)[None, :]], -1)
        else:
            want = np.full((M, N), O.INF)
            wantp = np.full((M, N), -1)
        np.testing.assert_array_equal(Cm, want)
        np.testing.assert_array_equal(Cp, wantp)
        assert (Cp[0] == -1).all()


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build absent")
def test_oracle_min_plus_path_vs_reference():
    rng = np.random.default_rng(4)
    A = rng.integers(0, 5, (33, 20)).astype(np.int32)
    B = rng.integers(0, 5, (20, 41)).astype(np.int32)
    A[rng.random(A.shape) < 0.2] = O.INF
    B[rng.random(B.shape) < 0.2] = O.INF
    Bp = rng.integers(0, 10000, B.shape).astype(np.int32)
    Cm, Cp = O.min_plus(A, B, Bp)
    rc, rp = O.ref_min_plus_path(A, B, Bp)
    np.testing.assert_array_equal(Cm, O.from_ref(rc))
    np.testing.assert_array_equal(Cp, rp)


def test_fw_hop_check_accepts_floyd_paths_and_flags_mistakes():
    """The FW hop checker accepts the reference-semantics FW's own hops
    (orc_floyd with identity labels: path[i][j] = path[k][j]) on matrices
    with zero-weight entries, and flags planted mistakes."""
    rng = np.random.default_rng(5)
    n = 60
    m = rng.integers(0, 6, (n, n)).astype(np.int32)
    m[rng.random((n, n)) < 0.85] = O.INF
    m[:, n - 1] = O.INF                        # vertex n-1 unreachable from the others
    np.fill_diagonal(m, 0)
    lab = np.where(m < O.INF, np.arange(n, dtype=np.int32)[:, None], -1).astype(np.int32)
    np.fill_diagonal(lab, np.arange(n))
    d, p = O.floyd(m, with_path=lab)
    assert (m == 0).sum() > n                  # zero-weight arcs present
    assert O.fw_hop_check(m, d, p) == 0
    bad = p.copy()
    i, j = np.argwhere((d < O.INF) & (d < m))[0]
    bad[i, j] = (bad[i, j] + 1) % n            # a hop that is not tight (or not an arc)
    assert O.fw_hop_check(m, d, bad) > 0
    bad = p.copy()
    i, j = np.argwhere(d >= O.INF)[0]
    bad[i, j] = i                              # an unreachable cell with a hop
    assert O.fw_hop_check(m, d, bad) > 0


def test_min_plus_acc_vs_numpy():
    rng = np.random.default_rng(5)
    A = rng.integers(0, 50, (9, 11)).astype(np.int32)
    B = rng.integers(0, 50, (11, 6)).astype(np.int32)
    Cin = rng.integers(0, 80, (9, 6)).astype(np.int32)
    want = np.minimum(Cin, (A[:, :, None] + B[None, :, :]).min(1))
    np.testing.assert_array_equal(O.min_plus_acc(A, B, Cin), want)


# ---------- glue: closed-form partition ----------

def halves_partition(R, C):
    """R x C unit grid cut into left (cols < C/2) and right halves."""
    u, v, w = synth.grid_graph(R, C, 0.0, 0.0, 1, 1, 0)
    ro, col, wt = synth.to_csr(R * C, u, v, w, False)
    gid = np.where(np.arange(R * C) % C < C // 2, 1, 2).astype(np.int32)
    return ro, col, wt, gid


def test_boundary_of_halves_is_the_cut():
    R, C = 6, 8
    ro, col, wt, gid = halves_partition(R, C)
    b = O.find_boundary(2, gid, ro, col)
    c = np.arange(R * C) % C
    np.testing.assert_array_equal(b, ((c == C // 2 - 1) | (c == C // 2)).astype(np.int32))
    pc = O.sort_and_encode(2, gid, b)
    assert list(pc["ver_num"][1:]) == [R * C // 2] * 2
    assert list(pc["bdy_num"][1:]) == [R, R]
    assert list(pc["ver_off"][1:]) == [0, R * C // 2]
    st = pc["st2ed"]
    # boundary first inside each part, parts in order
    assert (gid[st[: R * C // 2]] == 1).all() and (b[st[:R]] == 1).all()
    assert (b[st[R: R * C // 2]] == 0).all()
    np.testing.assert_array_equal(pc["ed2st"][st], np.arange(R * C))


def test_submat_build_and_decode_on_halves():
    R, C = 6, 8
    ro, col, wt, gid = halves_partition(R, C)
    pc = O.precondition(2, gid, ro, col)
    V = R * C
    for pid in (1, 2):
        n, b, s = int(pc["ver_num"][pid]), int(pc["bdy_num"][pid]), int(pc["ver_off"][pid])
        verts = pc["st2ed"][s:s + n]
        rb = O.batched_sssp(verts[:b], ro, col, wt)
        sub = O.submat_build(rb, gid, s, n, b, pc["st2ed"], pc["ed2st"], ro, col, wt)
        # boundary rows are the global distances among the part's vertices
        np.testing.assert_array_equal(sub[:b], rb[:, verts])
        # inner rows: 0 on the diagonal, 1 to grid neighbours in the part, else INF
        r, c = np.divmod(verts, C)
        adj = (np.abs(r[:, None] - r[None, :]) + np.abs(c[:, None] - c[None, :])) == 1
        want = np.where(adj, 1, INF)
        np.fill_diagonal(want, 0)
        np.testing.assert_array_equal(sub[b:], want[b:])
        full = np.zeros((n, V), np.int32)
        dec = O.submat_decode(sub, full, s, pc["st2ed"])
        np.testing.assert_array_equal(dec[:, verts], sub)


# ---------- partition invariance of the whole pipeline ----------

@pytest.mark.parametrize("K", [1, 2, 5])
def test_run_part_equals_dijkstra_rows(K):
    u, v, w = synth.grid_graph(12, 12, 0.1, 0.05, 1, 30, 7)
    V = 144
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    gid = (np.arange(V) * K // V + 1).astype(np.int32)  # contiguous id ranges
    pc = O.precondition(K, gid, ro, col)
    for pid in range(1, K + 1):
        s, n = int(pc["ver_off"][pid]), int(pc["ver_num"][pid])
        blk = O.run_part(pid, ro, col, wt, gid, pc)
        np.testing.assert_array_equal(blk, O.batched_sssp(pc["st2ed"][s:s + n], ro, col, wt))


# ---------- against the reference's own CPU code ----------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@needs_ref
def test_ref_maxvalue():
    assert O.ref().ref_maxvalue() == np.float32(1e8)


@needs_ref
@pytest.mark.parametrize("name", ["grid8.npz", "rgg100.npz"])
def test_oracle_dijkstra_vs_reference(name):
    g = golden(name)
    ro, col, w = g["row_offset"], g["col"], g["wt"]
    for s in range(0, len(ro) - 1, 7):
        d, _ = O.ref_dijkstra(s, ro, col, w)
        np.testing.assert_array_equal(O.dijkstra(s, ro, col, w), O.from_ref(d))


@needs_ref
def test_oracle_floyd_and_minplus_vs_reference():
    d = golden("dense70.npz")
    np.testing.assert_array_equal(O.floyd(d["mat"]), O.from_ref(O.ref_floyd(d["mat"])))
    m = golden("minplus_13x37x29.npz")
    np.testing.assert_array_equal(O.min_plus(m["A"], m["B"]),
                                  O.from_ref(O.ref_min_plus(m["A"], m["B"])))


@needs_ref
@pytest.mark.parametrize("K", [3, 8])
def test_oracle_precondition_vs_reference(K):
    u, v, w = synth.grid_graph(20, 20, 0.1, 0.02, 1, 9, 21)
    V = 400
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    rng = np.random.default_rng(K)
    gid = rng.integers(1, K + 1, V).astype(np.int32)
    mine = O.precondition(K, gid, ro, col)
    ref = O.ref_precondition(K, gid, ro, col)
    for key in ("is_bdy", "st2ed", "ed2st", "ver_num", "ver_off", "bdy_num", "bdy_off"):
        np.testing.assert_array_equal(mine[key][1:] if key.startswith(("ver", "bdy")) else mine[key],
                                      ref[key][1:] if key.startswith(("ver", "bdy")) else ref[key],
                                      err_msg=key)


@needs_ref
def test_oracle_run_part_vs_reference():
    u, v, w = synth.grid_graph(16, 16, 0.05, 0.02, 1, 40, 31)
    V = 256
    ro, col, wt = synth.to_csr(V, u, v, w, False)
    K = 4
    gid = (np.arange(V) * K // V + 1).astype(np.int32)
    pc = O.precondition(K, gid, ro, col)
    for pid in (1, 3):
        np.testing.assert_array_equal(O.run_part(pid, ro, col, wt, gid, pc),
                                      O.from_ref(O.ref_run_part(pid, ro, col, wt, gid, pc)))


# ---------------- predecessor rows (path matrices) ----------------

def _pred_graph(seed, zero=False):
    from fastapsp_amd import synth
    au, av, aw = synth.rgg_graph(200, 5, 0 if zero else 1, 9, seed)
    ro, col, w = synth.to_csr(200, au, av, aw, True)
    return ro, col, w


@pytest.mark.parametrize("zero", [False, True])
def test_pred_checker_accepts_dijkstra_parents(zero):
    """The checker passes the oracle's Dijkstra parent rows (the reference's
    dijkstra_path semantics, fap/kernel/batch_sssp.h:82-83,110), zero-weight
    arcs included."""
    ro, col, w = _pred_graph(3, zero)
    own = np.arange(0, 200, 7, dtype=np.int32)
    D, P = [], []
    for s in own:
        d, p = O.dijkstra(int(s), ro, col, w, with_path=True)
        D.append(d)
        P.append(p)
    D, P = np.array(D), np.array(P)
    assert O.check_pred(D, P, own, ro, col, w) == 0
    if O.ref_available():
        for i, s in enumerate(own[:5]):
            rd, rp = O.ref_dijkstra(int(s), ro, col, w)
            np.testing.assert_array_equal(O.from_ref(rd), D[i])
            assert O.check_pred(D[i:i + 1], rp[None, :], own[i:i + 1], ro, col, w) == 0


def test_pred_checker_accepts_pipeline_paths():
    """Path rows of the oracle's whole per-part pipeline (Dijkstra, FW and
    min-plus path updates) and, where built, of the reference's own run()."""
    from fastapsp_amd import synth
    u, v, ww = synth.grid_graph(20, 20, 0.1, 0.03, 1, 30, 4)
    V = 400
    ro, col, w = synth.to_csr(V, u, v, ww, False)
    gid = (np.arange(V) * 4 // V + 1).astype(np.int32)
    pc = O.precondition(4, gid, ro, col)
    for pid in (1, 3):
        d, p = O.run_part(pid, ro, col, w, gid, pc, with_path=True)
        s = int(pc["ver_off"][pid])
        own = pc["st2ed"][s:s + int(pc["ver_num"][pid])]
        assert O.check_pred(d, p, own, ro, col, w) == 0
        if O.ref_available():
            rd, rp = O.ref_run_part(pid, ro, col, w, gid, pc, with_path=True)
            assert O.check_pred(O.from_ref(rd), rp, own, ro, col, w) == 0


def test_pred_checker_flags_planted_mistakes():
    ro, col, w = _pred_graph(5, zero=True)
    own = np.array([0, 11], np.int32)
    D = np.array([O.dijkstra(int(s), ro, col, w) for s in own])
    P = np.array([O.dijkstra(int(s), ro, col, w, with_path=True)[1] for s in own])
    assert O.check_pred(D, P, own, ro, col, w) == 0
    reach = np.flatnonzero((D[0] < O.INF) & (np.arange(200) != own[0]))
    j = int(reach[len(reach) // 2])
    bad = P.copy()
    bad[0, own[0]] = -1                                   # own vertex not itself
    assert O.check_pred(D, bad, own, ro, col, w) > 0
    bad = P.copy()
    bad[0, j] = -1                                        # reachable marked unreachable
    assert O.check_pred(D, bad, own, ro, col, w) > 0
    bad = P.copy()
    nontight = [x for x in range(200) if x != P[0, j] and x != j]
    bad[0, j] = nontight[0]                               # not a tight in-neighbour
    assert O.check_pred(D, bad, own, ro, col, w) > 0
    unreach = np.flatnonzero(D[0] >= O.INF)
    if unreach.size:
        bad = P.copy()
        bad[0, unreach[0]] = own[0]                       # unreachable given a predecessor
        assert O.check_pred(D, bad, own, ro, col, w) > 0
    # a cycle through a zero-weight plateau: 0 -> 1, 0 -> 2 (w 5), 1 <-> 2 (w 0)
    ro2 = np.array([0, 2, 3, 4], np.int32)
    col2 = np.array([1, 2, 2, 1], np.int32)
    w2 = np.array([5, 5, 0, 0], np.int32)
    D2 = np.array([[0, 5, 5]], np.int32)
    assert O.check_pred(D2, np.array([[0, 0, 0]]), [0], ro2, col2, w2) == 0
    assert O.check_pred(D2, np.array([[0, 2, 0]]), [0], ro2, col2, w2) == 0
    assert O.check_pred(D2, np.array([[0, 2, 1]]), [0], ro2, col2, w2) > 0
