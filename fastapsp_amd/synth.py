"""Seeded synthetic inputs for the subgraph-solve path (numpy only).

Every random draw comes from a counter-based splitmix64 stream, so a graph is a
pure function of (generator, sizes, seed) and is identical on every machine and
numpy version: value i of stream (seed, tag) is splitmix64(key + i * GOLDEN)
with key = splitmix64(seed ^ (tag << 32)).  Integer ranges use the high 32 bits
times the range width (Lemire's multiply-shift), floats use the top 53 bits.

Generators (all simple graphs: no self loops, no duplicate arcs):

* ``grid_graph``   -- rows x cols 4-neighbour grid, a fraction of edges removed
  uniformly at random, optional random diagonal shortcuts, weights in [lo, hi].
  64x64 / 5% removed / unit weights is the small plain-method workload;
  512x512 / 10% removed / 2% of V diagonals / weights 1..100 the large one.
* ``rgg_graph``    -- random geometric graph: points uniform in the unit square,
  radius sqrt(deg / (pi n)), both arcs of every pair with independent weights.
* ``dense_matrix`` -- n x n distance matrix, each off-diagonal entry an arc
  with probability p and weight in [lo, hi], otherwise INF; diagonal 0.

Undirected graphs are returned as edge lists (u, v, w) with u < v or u > v as
generated; ``to_csr`` expands them to arcs.  ``write_mtx`` writes the
reference's input format (header "n n m", 1-based "a b [w]" lines), which the
reference reads head-first so that each adjacency list comes out in reverse
file order (fap/utils/readMatFile.h:30-34, 147-162); ``csr_reference_order``
reproduces that order.
"""

from __future__ import annotations

import math
import os

import numpy as np

INF = 0x3FFFFFFF
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# stream tags
T_DROP, T_DIAG, T_WEIGHT, T_X, T_Y, T_EDGE, T_SIZE, T_ARRIVAL, T_KIND, T_QUERY = range(1, 11)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def splitmix64(x) -> np.ndarray:
    """splitmix64 output for state x (the state is advanced by GOLDEN first)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(x + GOLDEN)


def stream(seed: int, tag: int, n: int, offset: int = 0) -> np.ndarray:
    """Values offset .. offset+n-1 of stream (seed, tag) as uint64."""
    key = splitmix64(np.uint64((seed & 0xFFFFFFFFFFFFFFFF) ^ ((tag & 0xFFFFFFFF) << 32)))
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(key + idx * GOLDEN + GOLDEN)


def uniform_int(u: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """Map uint64 draws to integers in [lo, hi] (hi - lo + 1 <= 2**32)."""
    width = np.uint64(hi - lo + 1)
    return (np.int64(lo) + ((u >> np.uint64(32)) * width >> np.uint64(32)).astype(np.int64))


def uniform01(u: np.ndarray) -> np.ndarray:
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _pick_lowest(keys: np.ndarray, m: int) -> np.ndarray:
    """Boolean mask of the m entries with the smallest keys (ties by index)."""
    order = np.lexsort((np.arange(keys.size), keys))
    mask = np.zeros(keys.size, dtype=bool)
    mask[order[:m]] = True
    return mask


def grid_graph(rows: int, cols: int, drop_frac: float, diag_frac: float,
               wlo: int, whi: int, seed: int):
    """Undirected grid edge list (u, v, w) as three int32 arrays, V = rows*cols.

    drop_frac of the 4-neighbour edges are removed (round(drop_frac * E4)
    edges, chosen uniformly); round(diag_frac * V) diagonal shortcuts
    (r,c)-(r+1,c+1) or (r,c+1)-(r+1,c) are added, chosen uniformly among all
    2*(rows-1)*(cols-1) candidates; weights uniform in [wlo, whi].
    """
    r = np.arange(rows, dtype=np.int64)[:, None]
    c = np.arange(cols, dtype=np.int64)[None, :]
    vid = r * cols + c
    hu, hv = vid[:, :-1].ravel(), vid[:, 1:].ravel()
    vu, vv = vid[:-1, :].ravel(), vid[1:, :].ravel()
    eu = np.concatenate([hu, vu])
    ev = np.concatenate([hv, vv])
    n_drop = int(round(drop_frac * eu.size))
    drop = _pick_lowest(stream(seed, T_DROP, eu.size), n_drop)
    eu, ev = eu[~drop], ev[~drop]
    n_diag = int(round(diag_frac * rows * cols))
    if n_diag > 0:
        d1u, d1v = vid[:-1, :-1].ravel(), vid[1:, 1:].ravel()
        d2u, d2v = vid[:-1, 1:].ravel(), vid[1:, :-1].ravel()
        du = np.concatenate([d1u, d2u])
        dv = np.concatenate([d1v, d2v])
        pick = _pick_lowest(stream(seed, T_DIAG, du.size), n_diag)
        eu = np.concatenate([eu, du[pick]])
        ev = np.concatenate([ev, dv[pick]])
    if wlo == whi:
        w = np.full(eu.size, wlo, dtype=np.int64)
    else:
        w = uniform_int(stream(seed, T_WEIGHT, eu.size), wlo, whi)
    return eu.astype(np.int32), ev.astype(np.int32), w.astype(np.int32)


def rgg_graph(n: int, avg_degree: float, wlo: int, whi: int, seed: int):
    """Directed random geometric graph as arc arrays (u, v, w), both arcs of
    every pair within radius sqrt(avg_degree / (pi n)), independent weights."""
    x = uniform01(stream(seed, T_X, n))
    y = uniform01(stream(seed, T_Y, n))
    rad = math.sqrt(avg_degree / (math.pi * n))
    from scipy.spatial import cKDTree  # deterministic pair set within rad
    pairs = cKDTree(np.stack([x, y], 1)).query_pairs(rad, output_type="ndarray")
    pu = pairs[:, 0].astype(np.int64)
    pv = pairs[:, 1].astype(np.int64)
    lo_ = np.minimum(pu, pv)
    hi_ = np.maximum(pu, pv)
    o = np.lexsort((hi_, lo_))
    lo_, hi_ = lo_[o], hi_[o]
    au = np.concatenate([lo_, hi_])
    av = np.concatenate([hi_, lo_])
    w = uniform_int(stream(seed, T_WEIGHT, au.size), wlo, whi)
    return au.astype(np.int32), av.astype(np.int32), w.astype(np.int32)


def to_csr(V: int, u: np.ndarray, v: np.ndarray, w: np.ndarray, directed: bool):
    """CSR (row_offset[V+1], col, w) with arcs grouped by source in input order.
    Undirected edge lists get both arcs."""
    if not directed:
        u, v, w = np.concatenate([u, v]), np.concatenate([v, u]), np.concatenate([w, w])
    order = np.argsort(u, kind="stable")
    col = v[order].astype(np.int32)
    wt = w[order].astype(np.int32)
    counts = np.bincount(u, minlength=V)
    ro = np.zeros(V + 1, dtype=np.int32)
    np.cumsum(counts, out=ro[1:])
    return ro, col, wt


def csr_reference_order(V: int, u, v, w, directed: bool):
    """CSR in the order the reference's reader produces from a file listing
    edges (u[i]+1, v[i]+1, w[i]) line by line: arcs pushed at the head of
    per-vertex lists (a->b, then b->a if undirected), so each adjacency list
    is the reverse of insertion order (fap/utils/readMatFile.h:30-34,112-162).
    """
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.asarray(w, np.int64)
    if directed:
        au, av, aw = u, v, w
    else:
        au = np.stack([u, v], 1).ravel()
        av = np.stack([v, u], 1).ravel()
        aw = np.stack([w, w], 1).ravel()
    seq = np.arange(au.size)
    order = np.lexsort((-seq, au))
    col = av[order].astype(np.int32)
    wt = aw[order].astype(np.int32)
    counts = np.bincount(au, minlength=V)
    ro = np.zeros(V + 1, dtype=np.int32)
    np.cumsum(counts, out=ro[1:])
    return ro, col, wt


def write_mtx(path: str, V: int, u, v, w=None):
    """Write the reference's .mtx flavour: "V V m" then 1-based "a b [w]"."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    u = np.asarray(u, np.int64) + 1
    v = np.asarray(v, np.int64) + 1
    with open(path, "w") as f:
        f.write(f"{V} {V} {u.size}\n")
        if w is None:
            np.savetxt(f, np.stack([u, v], 1), fmt="%d")
        else:
            np.savetxt(f, np.stack([u, v, np.asarray(w, np.int64)], 1), fmt="%d")


def dense_matrix(n: int, p: float, wlo: int, whi: int, seed: int,
                 rows: slice | None = None) -> np.ndarray:
    """n x n int32 matrix (or the row range `rows`): arc with probability p and
    weight in [wlo, whi], else INF; diagonal 0."""
    r0 = 0 if rows is None else rows.start
    r1 = n if rows is None else rows.stop
    out = np.empty((r1 - r0, n), dtype=np.int32)
    chunk = max(1, (1 << 22) // n)
    for a in range(r0, r1, chunk):
        b = min(r1, a + chunk)
        cnt = (b - a) * n
        ue = uniform01(stream(seed, T_EDGE, cnt, a * n))
        uw = uniform_int(stream(seed, T_WEIGHT, cnt, a * n), wlo, whi)
        blk = np.where(ue < p, uw, INF).astype(np.int32).reshape(b - a, n)
        ii = np.arange(a, b)
        blk[ii - a, ii] = 0
        out[a - r0:b - r0] = blk
    return out


def loguniform_sizes(count: int, lo: int, hi: int, seed: int, tag: int = T_SIZE) -> np.ndarray:
    """count integers log-uniform in [lo, hi]."""
    u = uniform01(stream(seed, tag, count))
    s = np.exp(math.log(lo) + u * (math.log(hi + 1) - math.log(lo)))
    return np.clip(np.floor(s).astype(np.int64), lo, hi)
