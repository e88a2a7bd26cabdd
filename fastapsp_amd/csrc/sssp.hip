// sssp.hip -- multi-source SSSP (stage 1): one workgroup per source,
// near-far frontier processing, no grid-wide synchronisation.
//
// Replaces handle_boundry_path (fap/kernel/batch_sssp.h:254-277), whose CPU
// branch runs one binary-heap Dijkstra per boundary vertex (:79-114, OpenMP
// over sources :232-252) and whose HIP branch runs the sources one after
// another with a host round trip per Bellman-Ford round
// (fap/kernel/hip/batch_sssp.cpp:91-158).  Shortest distances are unique, so
// any exact method reproduces the reference's rows bit for bit.
//
// Method (near-far, Davidson et al. 2014, restated): each workgroup owns one
// source at a time and keeps, in global memory, two "near" queues (current
// and next round) and two "far" piles.  A round relaxes the out-arcs of every
// near vertex u (strict: atomicMin(dist[v], dist[u] + w) only when a plain
// read shows the candidate is smaller).  An improved v goes to the next near
// queue if its new distance is below the threshold, otherwise to the far pile.
// When the near queue drains, the threshold moves to min(far distance) + delta
// and the far pile is split.  Vertices are therefore settled in near
// distance order, as Dijkstra's buckets would, and re-relaxations are limited
// to a band of width delta.
//
// Queue appends are compacted per wavefront: a ballot of the pushing lanes,
// one LDS atomicAdd per wave, lane slot = mbcnt of the ballot.  Duplicates
// are suppressed by a per-vertex round stamp (near) and a per-source tag
// (far); both arrays are never cleared, the stamps only grow.
//
// Every loop is bounded: a vertex enters a queue only after a strict
// improvement of its distance, and each far split raises the threshold above
// the smallest far distance.  Workgroups never wait on each other.
#include <algorithm>
#include <vector>

#include "common.h"

namespace fapk {


struct NfState {
  uint32_t *qa, *qb, *fa, *fb;  // near current/next, far current/next (V each)
  int32_t *stamp;               // near round stamp (V)
  int32_t *farmark;             // far-pile tag (V)
};

__device__ __forceinline__ unsigned lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Append `item` for lanes with `push`; `cnt` is an LDS counter.
__device__ __forceinline__ void wave_push(bool push, uint32_t item, uint32_t *q, uint32_t *cnt) {
  const unsigned long long mask = __ballot(push);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (push) q[base + lane_prefix(mask)] = item;
}

// Reorders queue `q` (n entries) by vertex block (v >> bsh) into `tmp`:
// counting sort over NBINS LDS bins, entries of one bin in arbitrary order.
// Consecutive lanes then take vertices of the same block, so their distance
// reads, row offsets and their neighbours' distances share cache lines.
// Called by the whole workgroup after a barrier; ends with a barrier.
constexpr int NBINS = 4096;

template <int NF_THREADS>
__device__ __forceinline__ void bin_queue(const uint32_t *__restrict__ q, uint32_t *__restrict__ tmp,
                                          uint32_t n, int bsh, uint32_t *hist, int *wsum) {
  constexpr int PER = NBINS / NF_THREADS;
  constexpr int NW = NF_THREADS / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int b = tid; b < NBINS; b += NF_THREADS) hist[b] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < n; i += NF_THREADS) atomicAdd(&hist[q[i] >> bsh], 1u);
  __syncthreads();
  uint32_t loc[PER];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    loc[k] = hist[tid * PER + k];
    sum += loc[k];
  }
  uint32_t x = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = (int)x;
  __syncthreads();
  if (wave == 0) {
    uint32_t t = lane < NW ? (uint32_t)wsum[lane] : 0u;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)t, off);
      if (lane >= off) t += y;
    }
    if (lane < NW) wsum[lane] = (int)t;
  }
  __syncthreads();
  uint32_t base = (wave > 0 ? (uint32_t)wsum[wave - 1] : 0u) + x - sum;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    hist[tid * PER + k] = base;
    base += loc[k];
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += NF_THREADS) {
    const uint32_t v = q[i];
    tmp[atomicAdd(&hist[v >> bsh], 1u)] = v;
  }
  __syncthreads();
}

template <int NF_THREADS, bool BIN>
__global__ __launch_bounds__(NF_THREADS) void sssp_nf_kernel(
    const int32_t *__restrict__ row_offset, const int2 *__restrict__ edges, int V,
    const int32_t *__restrict__ sources, int S, int32_t *out, int64_t ld,
    int32_t *const *__restrict__ rowptr, uint32_t *ws, int64_t ws_stride, int delta, int bsh,
    uint32_t bin_min) {
  // Queue counters live in LDS in three sets of {next near, new far}.  Each
  // phase (a near round or a far split) appends into set ci, thread 0 zeroes
  // set ci+1 (mod 3) -- last read before the previous barrier, next written
  // after this one -- then one barrier, then every thread reads set ci.
  __shared__ uint32_t s_cnt[6];
  __shared__ int s_min;
  __shared__ uint32_t s_hist[BIN ? NBINS : 1];
  __shared__ int s_wsum[BIN ? NF_THREADS / 64 : 1];
  uint32_t *base = ws + (int64_t)blockIdx.x * ws_stride;
  NfState st;
  st.qa = base;
  st.qb = base + V;
  st.fa = base + 2 * (int64_t)V;
  st.fb = base + 3 * (int64_t)V;
  st.stamp = (int32_t *)(base + 4 * (int64_t)V);
  st.farmark = (int32_t *)(base + 5 * (int64_t)V);
  const int tid = threadIdx.x;
  // stamps are per workgroup and only grow: start from 0 once per launch
  for (int v = tid; v < V; v += NF_THREADS) {
    st.stamp[v] = 0;
    st.farmark[v] = -1;
  }
  if (tid < 6) s_cnt[tid] = 0;
  int round_tag = 0;
  int ci = 0;
  for (int si = blockIdx.x; si < S; si += gridDim.x) {
    int32_t *dist = rowptr ? rowptr[si] : out + (int64_t)si * ld;
    const int src = sources[si];
    for (int v = tid; v < V; v += NF_THREADS) dist[v] = (v == src) ? 0 : 0x3FFFFFFF;
    uint32_t *qa = st.qa, *qb = st.qb, *fa = st.fa, *fb = st.fb;
    uint32_t n_near = 1, n_far = 0;
    if (tid == 0) qa[0] = (uint32_t)src;
    int thr = delta;
    const int tag = si;  // distinct per source handled by this workgroup
    __syncthreads();
    for (;;) {
      // ---- drain the near queue, one barrier per round ----
      while (n_near > 0) {
        round_tag++;
        uint32_t *cn = &s_cnt[2 * ci];
        for (uint32_t i0 = 0; i0 < n_near; i0 += NF_THREADS) {
          const uint32_t i = i0 + tid;
          const bool valid = i < n_near;
          int du = 0, e = 0, e_end = 0;
          if (valid) {
            const int u = (int)qa[i];
            du = dist[u];
            e = row_offset[u];
            e_end = row_offset[u + 1];
          }
          while (__any(valid && e < e_end)) {
            bool pn = false, pf = false;
            uint32_t item = 0;
            if (valid && e < e_end) {
              const int2 ed = edges[e];
              e++;
              const int nd = du + ed.y;
              int32_t *dp = dist + ed.x;
              if (nd < *dp) {
                const int old = atomicMin(dp, nd);
                if (nd < old) {
                  item = (uint32_t)ed.x;
                  if (nd < thr) {
                    pn = atomicMax(st.stamp + ed.x, round_tag) < round_tag;
                  } else {
                    pf = atomicExch(st.farmark + ed.x, tag) != tag;
                  }
                }
              }
            }
            wave_push(pn, item, qb, &cn[0]);
            wave_push(pf, item, fa + n_far, &cn[1]);
          }
        }
        const int cz = ci == 2 ? 0 : ci + 1;
        if (tid == 0) {
          s_cnt[2 * cz] = 0;
          s_cnt[2 * cz + 1] = 0;
        }
        __syncthreads();
        n_near = cn[0];
        n_far += cn[1];
        ci = cz;
        uint32_t *t = qa;
        qa = qb;
        qb = t;
        if (BIN && n_near >= bin_min) {
          bin_queue<NF_THREADS>(qa, qb, n_near, bsh, s_hist, s_wsum);
          t = qa;
          qa = qb;
          qb = t;
        }
      }
      if (n_far == 0) break;
      // ---- threshold = min far distance + delta ----
      if (tid == 0) s_min = 0x7FFFFFFF;
      __syncthreads();
      int lmin = 0x7FFFFFFF;
      for (uint32_t i = tid; i < n_far; i += NF_THREADS) {
        const int v = (int)fa[i];
        if (st.farmark[v] == tag) lmin = min(lmin, dist[v]);
      }
      for (int off = 32; off > 0; off >>= 1) lmin = min(lmin, __shfl_down(lmin, off));
      if ((tid & 63) == 0) atomicMin(&s_min, lmin);
      __syncthreads();
      const int fmin = s_min;
      if (fmin == 0x7FFFFFFF) break;  // every far entry was stale
      thr = fmin + delta;
      round_tag++;
      // ---- split the far pile: below thr -> near, else stays far ----
      uint32_t *cn = &s_cnt[2 * ci];
      for (uint32_t i0 = 0; i0 < n_far; i0 += NF_THREADS) {
        const uint32_t i = i0 + tid;
        bool pn = false, pf = false;
        uint32_t item = 0;
        if (i < n_far) {
          const int v = (int)fa[i];
          item = (uint32_t)v;
          if (st.farmark[v] == tag) {
            if (dist[v] < thr) {
              st.farmark[v] = -1;
              pn = atomicMax(st.stamp + v, round_tag) < round_tag;
            } else {
              pf = true;
            }
          }
        }
        wave_push(pn, item, qa, &cn[0]);
        wave_push(pf, item, fb, &cn[1]);
      }
      const int cz = ci == 2 ? 0 : ci + 1;
      if (tid == 0) {
        s_cnt[2 * cz] = 0;
        s_cnt[2 * cz + 1] = 0;
      }
      __syncthreads();
      n_near = cn[0];
      n_far = cn[1];
      ci = cz;
      uint32_t *t = fa;
      fa = fb;
      fb = t;
      if (BIN && n_near >= bin_min) {
        bin_queue<NF_THREADS>(qa, qb, n_near, bsh, s_hist, s_wsum);
        t = qa;
        qa = qb;
        qb = t;
      }
    }
    // far piles are reused by the next source: tags make stale marks harmless
  }
}


// ---------------------------------------------------------------------------
// Grouped sources: one workgroup runs near-far for S sources at once.
//
// Distances live interleaved, D[v * S + s], so the S values of a vertex are
// one 16- or 32-byte segment.  A queue entry is a vertex; S consecutive lanes
// handle it together, lane s for source s, walking the vertex's arcs in
// lockstep: one arc load serves all S sources and the S distance reads of a
// neighbour are one segment.  Per vertex two masks record which sources
// still have to be relaxed from it: `dirty` (improved since the vertex was
// last processed) and `pend` (improved to a value at or above that source's
// threshold while the vertex waits in the far pile).  Processing a vertex
// takes its dirty mask (atomic exchange) and relaxes only those sources.
// Only a far split clears pend bits (no relaxation runs beside it), so a far
// improvement is never lost to a concurrent processing of its vertex; a pend
// bit left over after its value was relaxed only lowers that source's next
// threshold.  Each source keeps
// its own threshold (a register of its lanes); when the near queue drains,
// every source with pending far values moves its threshold to its smallest
// pending value + delta and the far pile is split per source.
//
// Relaxation is label-correcting: whenever D[v][s] decreases, bit s of
// dirty[v] is set and v is queued (near or far), so v is processed again;
// at exit every arc satisfies D[v][s] <= D[u][s] + w, so the values are the
// exact distances whatever order the relaxations ran in.
template <int S, int NT>
__global__ __launch_bounds__(NT) void sssp_group_kernel(
    const int32_t *__restrict__ row_offset, const int2 *__restrict__ edges, int V,
    const int32_t *__restrict__ sources, int n_src, int32_t *out, int64_t ld,
    int32_t *const *__restrict__ rowptr, uint32_t *ws, int64_t ws_stride, int64_t d_off,
    int delta) {
  constexpr int NG = NT / S;  // queue entries per pass
  constexpr uint32_t SMASK = (1u << S) - 1;
  __shared__ uint32_t s_cnt[6];
  __shared__ int s_min[S];
  __shared__ int32_t *s_rows[S];
  __shared__ uint32_t s_n0;
  uint32_t *base = ws + (int64_t)blockIdx.x * ws_stride;
  uint32_t *const qa0 = base, *const qb0 = base + V, *const fa0 = base + 2 * (int64_t)V,
                  *const fb0 = base + 3 * (int64_t)V;
  int32_t *const stamp = (int32_t *)(base + 4 * (int64_t)V);
  int32_t *const farmark = (int32_t *)(base + 5 * (int64_t)V);
  uint32_t *const dirty = base + 6 * (int64_t)V;
  uint32_t *const pend = base + 7 * (int64_t)V;
  int32_t *const D = (int32_t *)(base + d_off);  // V * S, 16-byte aligned
  const int tid = threadIdx.x, lane = tid & 63, sl = tid % S, grp = tid / S;
  const int gl = lane - sl;  // first lane of this lane's group
  for (int v = tid; v < V; v += NT) {
    stamp[v] = 0;
    farmark[v] = -1;
  }
  if (tid < 6) s_cnt[tid] = 0;
  int round_tag = 0;
  int ci = 0;
  const int ngroups = (n_src + S - 1) / S;
  for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const int s0 = gi * S;
    const int ns = min(S, n_src - s0);
    const int tag = gi;
    {
      const int4 inf4 = make_int4(FAP_INF, FAP_INF, FAP_INF, FAP_INF);
      int4 *D4 = reinterpret_cast<int4 *>(D);
      for (int64_t x = tid; x < (int64_t)V * S / 4; x += NT) D4[x] = inf4;
    }
    for (int v = tid; v < V; v += NT) {
      dirty[v] = 0;
      pend[v] = 0;
    }
    if (tid < ns) s_rows[tid] = rowptr ? rowptr[s0 + tid] : out + (int64_t)(s0 + tid) * ld;
    round_tag++;  // stamps of the initial queue
    __syncthreads();
    if (tid == 0) {
      uint32_t n0 = 0;
      for (int t = 0; t < ns; t++) {
        const int src = sources[s0 + t];
        D[(int64_t)src * S + t] = 0;
        dirty[src] |= 1u << t;
        if (stamp[src] != round_tag) {
          stamp[src] = round_tag;
          qa0[n0++] = (uint32_t)src;
        }
      }
      s_n0 = n0;
    }
    __syncthreads();
    uint32_t *qa = qa0, *qb = qb0, *fa = fa0, *fb = fb0;
    uint32_t n_near = s_n0, n_far = 0;
    int thr = delta;  // this lane's source's threshold
    for (;;) {
      // ---- near rounds ----
      while (n_near > 0) {
        round_tag++;
        uint32_t *cn = &s_cnt[2 * ci];
        for (uint32_t i0 = 0; i0 < n_near; i0 += NG) {
          const uint32_t i = i0 + grp;
          const bool valid = i < n_near;
          int u = 0;
          uint32_t m = 0;
          if (valid) {
            u = (int)qa[i];
            if (sl == 0) m = atomicExch(dirty + u, 0u);
          }
          m = (uint32_t)__shfl((int)m, gl);
          const bool act = (m >> sl) & 1;
          int du = 0, e = 0, e_end = 0;
          // read at device scope, after the mask was taken: an improvement
          // whose dirty bit is in m completed its atomicMin before that bit
          // was set, so this sees it
          if (act)
            du = __hip_atomic_load(D + (int64_t)u * S + sl, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
          if (m) {
            e = row_offset[u];
            e_end = row_offset[u + 1];
          }
          while (__any(e < e_end)) {
            bool in = false, fr = false;
            int v = 0;
            if (e < e_end) {
              const int2 ed = edges[e];
              e++;
              v = ed.x;
              if (act) {
                const int nd = du + ed.y;
                int32_t *dp = D + (int64_t)v * S + sl;
                if (nd < *dp) {
                  const int old = atomicMin(dp, nd);
                  if (nd < old) {
                    in = nd < thr;
                    fr = !in;
                  }
                }
              }
            }
            const uint32_t gn = (uint32_t)(__ballot(in) >> gl) & SMASK;
            const uint32_t gf = (uint32_t)(__ballot(fr) >> gl) & SMASK;
            bool pn = false, pf = false;
            if (sl == 0 && (gn | gf)) {
              atomicOr(dirty + v, gn | gf);
              if (gf) {
                atomicOr(pend + v, gf);
                pf = atomicExch(farmark + v, tag) != tag;
              }
              if (gn) pn = atomicMax(stamp + v, round_tag) < round_tag;
            }
            wave_push(pn, (uint32_t)v, qb, &cn[0]);
            wave_push(pf, (uint32_t)v, fa + n_far, &cn[1]);
          }
        }
        const int cz = ci == 2 ? 0 : ci + 1;
        if (tid == 0) {
          s_cnt[2 * cz] = 0;
          s_cnt[2 * cz + 1] = 0;
        }
        __syncthreads();
        n_near = cn[0];
        n_far += cn[1];
        ci = cz;
        uint32_t *t = qa;
        qa = qb;
        qb = t;
      }
      if (n_far == 0) break;
      // ---- per-source smallest pending far value ----
      if (tid < S) s_min[tid] = 0x7FFFFFFF;
      __syncthreads();
      int lmin = 0x7FFFFFFF;
      for (uint32_t i = grp; i < n_far; i += NG) {
        const int v = (int)fa[i];
        if (farmark[v] == tag && ((pend[v] >> sl) & 1)) lmin = min(lmin, D[(int64_t)v * S + sl]);
      }
      for (int off = 32; off >= S; off >>= 1) lmin = min(lmin, __shfl_xor(lmin, off));
      if (lane < S) atomicMin(&s_min[lane], lmin);
      __syncthreads();
      bool any = false;
#pragma unroll
      for (int t = 0; t < S; t++) any |= s_min[t] != 0x7FFFFFFF;
      if (!any) break;  // every far entry was stale
      if (s_min[sl] != 0x7FFFFFFF) thr = s_min[sl] + delta;
      round_tag++;
      // ---- split: pending values below the new thresholds go near ----
      uint32_t *cn = &s_cnt[2 * ci];
      for (uint32_t i0 = 0; i0 < n_far; i0 += NG) {
        const uint32_t i = i0 + grp;
        int v = 0;
        bool live = false;
        uint32_t p = 0;
        if (i < n_far) {
          v = (int)fa[i];
          live = farmark[v] == tag;
          if (live) p = pend[v];
        }
        const bool go = ((p >> sl) & 1) && D[(int64_t)v * S + sl] < thr;
        const uint32_t gb = (uint32_t)(__ballot(go) >> gl) & SMASK;
        bool pn = false, pf = false;
        if (sl == 0 && live) {
          const uint32_t rest = p & ~gb;
          if (gb) {
            pend[v] = rest;
            pn = atomicMax(stamp + v, round_tag) < round_tag;
          }
          if (rest) pf = true;
          else farmark[v] = -1;
        }
        wave_push(pn, (uint32_t)v, qa, &cn[0]);
        wave_push(pf, (uint32_t)v, fb, &cn[1]);
      }
      const int cz = ci == 2 ? 0 : ci + 1;
      if (tid == 0) {
        s_cnt[2 * cz] = 0;
        s_cnt[2 * cz + 1] = 0;
      }
      __syncthreads();
      n_near = cn[0];
      n_far = cn[1];
      ci = cz;
      uint32_t *t = fa;
      fa = fb;
      fb = t;
    }
    // ---- rows out: D[v][s] -> row s, column v ----
    // the last values were written by atomics at the device cache: drop any
    // older copies from this CU's vector cache before reading them plainly
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int v = tid; v < V; v += NT) {
      int32_t d[S];
      const int4 *src4 = reinterpret_cast<const int4 *>(D + (int64_t)v * S);
#pragma unroll
      for (int q = 0; q < S / 4; q++) {
        const int4 x = src4[q];
        d[4 * q] = x.x;
        d[4 * q + 1] = x.y;
        d[4 * q + 2] = x.z;
        d[4 * q + 3] = x.w;
      }
#pragma unroll
      for (int t = 0; t < S; t++)
        if (t < ns) s_rows[t][v] = d[t];
    }
    __syncthreads();  // D and s_rows are rewritten by the next group
  }
}

}  // namespace fapk

namespace fapi {

// Threshold width: 8 mean arc weights, the best of 1..32 x mean measured on
// the 512 x 512 road-like grid (w 1..100) and the 64 x 64 unit grid
// (bench.py --workload ssspsweep).

// Rows go to d_out + i * ld_out, or to d_rowptr[i] (a device array of row
// pointers) when d_rowptr is not null.
namespace {

template <int S, int NT>
void launch_group(int grid, hipStream_t s, const fap_graph_dev *g, const int32_t *src, int n_src,
                  int32_t *out, int64_t ld, int32_t *const *rowptr, uint32_t *ws,
                  int64_t stride, int64_t d_off, int delta) {
  hipLaunchKernelGGL((fapk::sssp_group_kernel<S, NT>), dim3(grid), dim3(NT), 0, s, g->row_offset,
                     g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride, d_off, delta);
}

template <int NT>
void launch_single(int grid, hipStream_t s, const fap_graph_dev *g, const int32_t *src,
                   int n_src, int32_t *out, int64_t ld, int32_t *const *rowptr, uint32_t *ws,
                   int64_t stride, int delta, int bin) {
  // blocks of 2^bsh vertices, at least 64 (4 lines of distances), at most
  // NBINS blocks; queues of at least bin_min entries are binned
  int bsh = 6;
  while (((int64_t)g->V >> bsh) >= fapk::NBINS) bsh++;
  const uint32_t bin_min = bin >= 3 ? NT / 2 : bin == 2 ? NT : 2 * NT;
  if (bin)
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, true>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       delta, bsh, bin_min);
  else
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, false>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       delta, bsh, bin_min);
}

}  // namespace

// Launch shape of an SSSP over n_src sources: settings, per-workgroup state
// (6 V words for single sources, 8 V + S V for groups of S, rows of 64
// words) and grid (as many workgroups as source groups, up to 4 / 2 / 1 per
// CU for 256 / 512 / 1024 threads, and an 8 GiB state budget).
int sssp_layout(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src, int *threads,
                int *delta, int *group, int64_t *stride, int64_t *d_off) {
  sssp_settings(ctx, g, n_src, threads, delta, group);
  const int64_t vw = ((int64_t)g->V + 63) & ~(int64_t)63;
  *d_off = 8 * vw;
  *stride = *group == 1 ? 6 * vw : *d_off + (int64_t)*group * vw;
  const int64_t units = (n_src + *group - 1) / *group;
  const int64_t by_mem = std::max<int64_t>(1, ((int64_t)8 << 30) / (*stride * 4));
  const int per_cu = *threads == 256 ? 4 : *threads == 512 ? 2 : 1;
  return (int)std::max<int64_t>(
      1, std::min<int64_t>({units, (int64_t)ctx->num_cus * per_cu, by_mem}));
}

size_t sssp_ws_bytes(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src) {
  int threads, delta, group;
  int64_t stride, d_off;
  const int grid = sssp_layout(ctx, g, n_src, &threads, &delta, &group, &stride, &d_off);
  return (size_t)grid * stride * sizeof(uint32_t);
}

// Rows go to d_out + i * ld_out, or to d_rowptr[i] (a device array of row
// pointers) when d_rowptr is not null.  Sources are taken in groups of
// `group` (1 = one source per workgroup pass, 4 or 8 = grouped kernel);
// consecutive sources share a group, so callers list nearby sources
// together (a part's boundary vertices are).
int run_sssp_rows(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
                  int32_t *d_out, int64_t ld_out, int32_t *const *d_rowptr, hipStream_t s) {
  if (n_src <= 0) return FAP_OK;
  const int V = g->V;
  if (V <= 0) return fail(FAP_E_ARG, "sssp: empty graph");
  int threads, delta, group;
  int64_t stride, d_off;
  const int grid = sssp_layout(ctx, g, n_src, &threads, &delta, &group, &stride, &d_off);
  FAP_TRY(ensure_ws(ctx, ctx->ws_queue, (size_t)grid * stride * sizeof(uint32_t)));
  uint32_t *ws = (uint32_t *)ctx->ws_queue.ptr;
  const int32_t *src = d_sources;
  if (group == 1) {
    const int bin = ctx->opt_sssp_bin;
    if (threads == 256) launch_single<256>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin);
    else if (threads == 512) launch_single<512>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin);
    else launch_single<1024>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin);
  } else if (group == 4) {
    if (threads == 256) launch_group<4, 256>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
    else if (threads == 512) launch_group<4, 512>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
    else launch_group<4, 1024>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
  } else {
    if (threads == 256) launch_group<8, 256>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
    else if (threads == 512) launch_group<8, 512>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
    else launch_group<8, 1024>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, d_off, delta);
  }
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int run_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
             int32_t *d_out, int64_t ld_out, hipStream_t s) {
  return run_sssp_rows(ctx, g, d_sources, n_src, d_out, ld_out, nullptr, s);
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_sssp_multi(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources,
                              int32_t n_src, int32_t *d_out, int64_t ld_out, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (n_src < 0) return fail(FAP_E_ARG, "sssp: negative source count");
  if (n_src == 0) return FAP_OK;
  if (!d_sources || !d_out) return fail(FAP_E_ARG, "sssp: null pointer");
  if (ld_out < g->V) return fail(FAP_E_ARG, "sssp: ld_out < V");
  hipStream_t s = pick_stream(ctx, stream);
  // validate the sources on the host: an out-of-range id would index outside
  // the output rows
  std::vector<int32_t> h(n_src);
  FAP_HIP(hipMemcpyAsync(h.data(), d_sources, sizeof(int32_t) * n_src, hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n_src; i++)
    if (h[i] < 0 || h[i] >= g->V)
      return fail(FAP_E_ARG, "sssp: source " + std::to_string(i) + " out of range");
  return run_sssp(ctx, g, d_sources, n_src, d_out, ld_out, s);
}

extern "C" int fap_graph_set_delta(fap_graph_dev *g, int32_t delta) {
  if (!g || delta < 0) return fail(FAP_E_ARG, "bad graph/delta");
  g->delta = delta;
  return FAP_OK;
}
