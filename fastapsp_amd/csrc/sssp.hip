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
// near vertex u (strictly: dist[v] is lowered to dist[u] + w only when that is
// smaller).  An improved v goes to the next near queue if its new distance is
// below the threshold, otherwise to the far pile.  When the near queue
// drains, the threshold moves to min(far distance) + delta and the far pile
// is split.  Vertices are therefore settled in near distance order, as
// Dijkstra's buckets would, and re-relaxations are limited to a band of width
// delta.
//
// Queue appends are compacted per wavefront: a ballot of the pushing lanes,
// one LDS atomicAdd per wave, lane slot = mbcnt of the ballot.  Duplicates
// are suppressed by queue state kept beside each distance, in the same word,
// so an improvement is one compare-and-swap on one cache line (see the two
// kernels below for the two word layouts).
//
// Every loop is bounded: a vertex enters a queue only after a strict
// improvement of its distance, and each far split raises the threshold above
// the smallest far distance.  Workgroups never wait on each other.
#include <algorithm>
#include <vector>

#include "common.h"

namespace fapk {


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

// ---------------------------------------------------------------------------
// Packed state (the default): a vertex's distance and queue tag in one
// 64-bit word (distance high, tag low), changed together by one
// compare-and-swap.  A relaxation touches one line of per-source state, not
// three (distance, round stamp and far tag in separate arrays measured 1.4x
// slower), and an improvement costs one atomic instead of two.  The working
// row belongs to the workgroup; the finished distances are copied to the
// output row with coalesced stores.
//
// Tags come from a per-source counter (every source rewrites all words):
//   0       never reached
//   R       pushed onto the next near queue by near round R, or moved to the
//           near queue by the far split that handed out R
//   ft      live in the current far pile; each split gives the entries that
//           stay far a new ft, and entries whose word carries another tag
//           (near-improved since they were pushed) are dropped
// An improvement to nd pushes v to the next near queue when nd < threshold
// and the old tag is not this round's R, to the far pile when nd >= threshold
// and the old tag is not ft.  Distances only decrease and the threshold is
// fixed between splits, so a vertex is live in the far pile at most once and
// the split rewrites words with plain stores.  Each failed compare-and-swap
// means another lane lowered the distance, so the retry loop is bounded.
__device__ __forceinline__ unsigned long long nf_word(int d, uint32_t tag) {
  return ((unsigned long long)(uint32_t)d << 32) | tag;
}

__device__ __forceinline__ int nf_dist(unsigned long long w) { return (int)(w >> 32); }

template <int NF_THREADS, bool BIN, int AB>
__global__ __launch_bounds__(NF_THREADS) void sssp_nf_kernel(
    const int32_t *__restrict__ row_offset, const int2 *__restrict__ edges, int V,
    const int32_t *__restrict__ sources, int S, int32_t *out, int64_t ld,
    int32_t *const *__restrict__ rowptr, uint32_t *ws, int64_t ws_stride, int64_t q_off,
    int delta, int bsh, uint32_t bin_min) {
  __shared__ uint32_t s_cnt[6];
  __shared__ int s_min;
  __shared__ uint32_t s_hist[BIN ? NBINS : 1];
  __shared__ int s_wsum[BIN ? NF_THREADS / 64 : 1];
  uint32_t *base = ws + (int64_t)blockIdx.x * ws_stride;
  unsigned long long *W = (unsigned long long *)base;  // V words, 8-byte aligned
  uint32_t *const q0 = base + q_off;
  const int tid = threadIdx.x;
  if (tid < 6) s_cnt[tid] = 0;
  int ci = 0;
  for (int si = blockIdx.x; si < S; si += gridDim.x) {
    int32_t *dist = rowptr ? rowptr[si] : out + (int64_t)si * ld;
    const int src = sources[si];
    for (int v = tid; v < V; v += NF_THREADS)
      W[v] = v == src ? nf_word(0, 1u) : nf_word(FAP_INF, 0u);
    uint32_t *qa = q0, *qb = q0 + V, *fa = q0 + 2 * (int64_t)V, *fb = q0 + 3 * (int64_t)V;
    uint32_t n_near = 1, n_far = 0;
    if (tid == 0) qa[0] = (uint32_t)src;
    uint32_t tags = 2;  // last tag handed out: 1 = the source's, 2 = the first far pile's
    uint32_t ft = 2;
    int thr = delta;
    __syncthreads();
    for (;;) {
      while (n_near > 0) {
        const uint32_t R = ++tags;
        uint32_t *cn = &s_cnt[2 * ci];
        for (uint32_t i0 = 0; i0 < n_near; i0 += NF_THREADS) {
          const uint32_t i = i0 + tid;
          const bool valid = i < n_near;
          int du = 0, e = 0, e_end = 0;
          if (valid) {
            const int u = (int)qa[i];
            du = nf_dist(W[u]);
            e = row_offset[u];
            e_end = row_offset[u + 1];
          }
          // AB arcs per lane per step: their arc and state loads are issued
          // together, so one step waits for two memory round trips, not 2 AB
          // (AB = 2 measured 4% faster than 1 and 2% faster than 4)
          while (__any(valid && e < e_end)) {
            int2 ed[AB];
            unsigned long long cur[AB];
#pragma unroll
            for (int k = 0; k < AB; k++) {
              ed[k] = make_int2(0, 0);
              if (valid && e + k < e_end) ed[k] = edges[e + k];
            }
#pragma unroll
            for (int k = 0; k < AB; k++) {
              cur[k] = 0;
              if (valid && e + k < e_end) cur[k] = W[ed[k].x];
            }
#pragma unroll
            for (int k = 0; k < AB; k++) {
              bool pn = false, pf = false;
              if (valid && e + k < e_end) {
                const int nd = du + ed[k].y;
                unsigned long long *wp = W + ed[k].x;
                unsigned long long c = cur[k];
                const bool near = nd < thr;
                const unsigned long long nw = nf_word(nd, near ? R : ft);
                while (nd < nf_dist(c)) {
                  const unsigned long long prev = atomicCAS(wp, c, nw);
                  if (prev == c) {
                    const uint32_t old_tag = (uint32_t)c;
                    if (near) pn = old_tag != R;
                    else pf = old_tag != ft;
                    break;
                  }
                  c = prev;
                }
              }
              wave_push(pn, (uint32_t)ed[k].x, qb, &cn[0]);
              wave_push(pf, (uint32_t)ed[k].x, fa + n_far, &cn[1]);
            }
            e += AB;
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
      // ---- threshold = min live far distance + delta ----
      if (tid == 0) s_min = 0x7FFFFFFF;
      __syncthreads();
      int lmin = 0x7FFFFFFF;
      for (uint32_t i = tid; i < n_far; i += NF_THREADS) {
        const unsigned long long w = W[fa[i]];
        if ((uint32_t)w == ft) lmin = min(lmin, nf_dist(w));
      }
      for (int off = 32; off > 0; off >>= 1) lmin = min(lmin, __shfl_down(lmin, off));
      if ((tid & 63) == 0) atomicMin(&s_min, lmin);
      __syncthreads();
      const int fmin = s_min;
      if (fmin == 0x7FFFFFFF) break;  // every far entry was stale
      thr = fmin + delta;
      const uint32_t rs = ++tags, fn = ++tags;
      // ---- split the far pile: below thr -> near, else stays far ----
      uint32_t *cn = &s_cnt[2 * ci];
      for (uint32_t i0 = 0; i0 < n_far; i0 += NF_THREADS) {
        const uint32_t i = i0 + tid;
        bool pn = false, pf = false;
        uint32_t item = 0;
        if (i < n_far) {
          const int v = (int)fa[i];
          item = (uint32_t)v;
          const unsigned long long w = W[v];
          if ((uint32_t)w == ft) {
            const int d = nf_dist(w);
            pn = d < thr;
            pf = !pn;
            W[v] = nf_word(d, pn ? rs : fn);
          }
        }
        wave_push(pn, item, qa, &cn[0]);
        wave_push(pf, item, fb, &cn[1]);
      }
      ft = fn;
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
    // every exit follows a barrier, and each thread copies (and the next
    // source re-initialises) the same words it owns: v = tid mod threads
    for (int v = tid; v < V; v += NF_THREADS) dist[v] = nf_dist(W[v]);
  }
}

// ---------------------------------------------------------------------------
// In-row state: distances stay below INF = 2^30 - 1, so a vertex's distance
// and two state bits fit one 32-bit word, d << 2 | state, kept in the output
// row itself (converted to plain distances in place at the end).  Per-source
// state is 4 bytes per vertex, and the workspace holds only the queues.
//   NQ  in the next near queue: set by a near push, cleared when the vertex
//       is popped -- by the atomic and-not that also reads its distance, so
//       an improvement either lands before the pop (and is used by it) or
//       sees NQ clear and queues the vertex again
//   NF  live in the far pile: set by a far push, cleared by a near
//       improvement (the pile entry turns stale) or by the split that moves
//       the vertex to the near queue
// A vertex is in a queue at most once (NQ) and live in the far pile at most
// once (NF), so every queue and pile holds at most V entries.  When a near
// round ends every queued vertex has been popped, so NQ is clear everywhere
// and the split may set it with plain stores.
constexpr uint32_t ST_NQ = 1u, ST_NF = 2u;

// Rows reached through a row-pointer array are global memory; saying so lets
// the state atomics and loads use global (not flat) instructions.
typedef __attribute__((address_space(1))) uint32_t gu32;

template <int NF_THREADS, bool BIN>
__global__ __launch_bounds__(NF_THREADS) void sssp_nfs_kernel(
    const int32_t *__restrict__ row_offset, const int2 *__restrict__ edges, int V,
    const int32_t *__restrict__ sources, int S, int32_t *out, int64_t ld,
    int32_t *const *__restrict__ rowptr, uint32_t *ws, int64_t ws_stride, int delta, int bsh,
    uint32_t bin_min) {
  __shared__ uint32_t s_cnt[6];
  __shared__ int s_min;
  __shared__ uint32_t s_hist[BIN ? NBINS : 1];
  __shared__ int s_wsum[BIN ? NF_THREADS / 64 : 1];
  uint32_t *const q0 = ws + (int64_t)blockIdx.x * ws_stride;
  const int tid = threadIdx.x;
  if (tid < 6) s_cnt[tid] = 0;
  int ci = 0;
  for (int si = blockIdx.x; si < S; si += gridDim.x) {
    gu32 *W = (gu32 *)(rowptr ? rowptr[si] : out + (int64_t)si * ld);
    const int src = sources[si];
    for (int v = tid; v < V; v += NF_THREADS)
      W[v] = v == src ? ST_NQ : (uint32_t)FAP_INF << 2;
    uint32_t *qa = q0, *qb = q0 + V, *fa = q0 + 2 * (int64_t)V, *fb = q0 + 3 * (int64_t)V;
    uint32_t n_near = 1, n_far = 0;
    if (tid == 0) qa[0] = (uint32_t)src;
    int thr = delta;
    __syncthreads();
    for (;;) {
      while (n_near > 0) {
        uint32_t *cn = &s_cnt[2 * ci];
        for (uint32_t i0 = 0; i0 < n_near; i0 += NF_THREADS) {
          const uint32_t i = i0 + tid;
          const bool valid = i < n_near;
          int du = 0, e = 0, e_end = 0;
          if (valid) {
            const int u = (int)qa[i];
            du = (int)(__hip_atomic_fetch_and(W + u, ~ST_NQ, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT) >> 2);
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
              gu32 *wp = W + ed.x;
              uint32_t cur = *wp;
              while (nd < (int)(cur >> 2)) {
                const uint32_t st = cur & 3u;
                const bool near = nd < thr;
                const uint32_t nw = ((uint32_t)nd << 2) | (near ? ST_NQ : (ST_NF | (st & ST_NQ)));
                uint32_t seen = cur;  // replaced by the word found when the swap fails
                if (__hip_atomic_compare_exchange_strong(wp, &seen, nw, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                  item = (uint32_t)ed.x;
                  if (near) pn = !(st & ST_NQ);
                  else pf = !(st & ST_NF);
                  break;
                }
                cur = seen;
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
      // ---- threshold = min live far distance + delta ----
      if (tid == 0) s_min = 0x7FFFFFFF;
      __syncthreads();
      int lmin = 0x7FFFFFFF;
      for (uint32_t i = tid; i < n_far; i += NF_THREADS) {
        const uint32_t w = W[fa[i]];
        if (w & ST_NF) lmin = min(lmin, (int)(w >> 2));
      }
      for (int off = 32; off > 0; off >>= 1) lmin = min(lmin, __shfl_down(lmin, off));
      if ((tid & 63) == 0) atomicMin(&s_min, lmin);
      __syncthreads();
      const int fmin = s_min;
      if (fmin == 0x7FFFFFFF) break;  // every far entry was stale
      thr = fmin + delta;
      // ---- split the far pile: below thr -> near, else stays far ----
      uint32_t *cn = &s_cnt[2 * ci];
      for (uint32_t i0 = 0; i0 < n_far; i0 += NF_THREADS) {
        const uint32_t i = i0 + tid;
        bool pn = false, pf = false;
        uint32_t item = 0;
        if (i < n_far) {
          const int v = (int)fa[i];
          item = (uint32_t)v;
          const uint32_t w = W[v];
          if (w & ST_NF) {
            pn = (int)(w >> 2) < thr;
            pf = !pn;
            if (pn) W[v] = (w & ~3u) | ST_NQ;
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
    // every exit follows a barrier, with NQ and NF clear everywhere (no queue
    // or live far entry is left); shift the distances back in place, each
    // thread the words it initialised
    for (int v = tid; v < V; v += NF_THREADS) W[v] >>= 2;
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
                   int64_t stride, int delta, int bin, int pack, int64_t q_off, int arcs) {
  // blocks of 2^bsh vertices, at least 64 (4 lines of distances), at most
  // NBINS blocks; queues of at least bin_min entries are binned
  int bsh = 6;
  while (((int64_t)g->V >> bsh) >= fapk::NBINS) bsh++;
  const uint32_t bin_min = bin >= 3 ? NT / 2 : bin == 2 ? NT : 2 * NT;
  if (pack == 2 && bin)
    hipLaunchKernelGGL((fapk::sssp_nfs_kernel<NT, true>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       delta, bsh, bin_min);
  else if (pack == 2)
    hipLaunchKernelGGL((fapk::sssp_nfs_kernel<NT, false>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       delta, bsh, bin_min);
  else if (bin && arcs == 4)
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, true, 4>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       q_off, delta, bsh, bin_min);
  else if (bin && arcs == 2)
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, true, 2>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       q_off, delta, bsh, bin_min);
  else if (bin)
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, true, 1>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       q_off, delta, bsh, bin_min);
  else
    hipLaunchKernelGGL((fapk::sssp_nf_kernel<NT, false, 2>), dim3(grid), dim3(NT), 0, s,
                       g->row_offset, g->edges, g->V, src, n_src, out, ld, rowptr, ws, stride,
                       q_off, delta, bsh, bin_min);
}

}  // namespace

// Launch shape of an SSSP over n_src sources: settings, per-workgroup state
// (4 or 6 V words for single sources, 8 V + S V for groups of S, rows of 64
// words) and grid (as many workgroups as source groups, up to 4 / 2 / 1 per
// CU for 256 / 512 / 1024 threads unless option sssp_wgs sets the count, and
// an 8 GiB state budget).
int sssp_layout(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src, int *threads,
                int *delta, int *group, int64_t *stride, int64_t *d_off) {
  sssp_settings(ctx, g, n_src, threads, delta, group);
  const int64_t vw = ((int64_t)g->V + 63) & ~(int64_t)63;
  *d_off = 8 * vw;
  // single sources: four queues, plus V 8-byte state words unless the state
  // lives in the output rows (sssp_pack 2)
  const int pack = ctx ? ctx->opt_sssp_pack : 1;
  *stride = *group == 1 ? (pack == 2 ? 4 : 6) * vw : *d_off + (int64_t)*group * vw;
  const int64_t units = (n_src + *group - 1) / *group;
  const int64_t by_mem = std::max<int64_t>(1, ((int64_t)8 << 30) / (*stride * 4));
  const int per_cu = (ctx && ctx->opt_sssp_wgs > 0) ? ctx->opt_sssp_wgs : 1024 / *threads;
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
    // state words first (V of 8 bytes), then the four queues
    const int bin = ctx->opt_sssp_bin;
    const int pack = ctx->opt_sssp_pack;
    const int arcs = ctx->opt_sssp_arcs;
    const int64_t q_off = 2 * (((int64_t)V + 63) & ~(int64_t)63);
    if (threads == 256) launch_single<256>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin, pack, q_off, arcs);
    else if (threads == 512) launch_single<512>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin, pack, q_off, arcs);
    else launch_single<1024>(grid, s, g, src, n_src, d_out, ld_out, d_rowptr, ws, stride, delta, bin, pack, q_off, arcs);
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
