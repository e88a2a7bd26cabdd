// bindex.hip -- stage 1 of the batched part solve through a boundary index.
//
// The reference's stage 1 (fap/kernel/batch_sssp.h:254-277, called by
// fapGraph::run, fap/fap.cpp:105-110) runs one SSSP over the whole graph
// from every boundary vertex of a part and keeps the rows: d(s, v) for s in
// B_P, v in V.  Here the same rows come out of two exact steps:
//
//  1. boundary distances on the boundary graph H.  Its vertices are the
//     boundary vertices of all parts; its arcs are the cut arcs of G (both
//     ends of a cut arc are boundary vertices, fap/utils/util_preprocess.h:
//     43-46) and, inside each part Q, an arc c -> c' of length L_Q(c, c'),
//     the distance from c to c' inside Q alone.  Any path of G between
//     boundary vertices splits into stretches inside one part, each from a
//     boundary vertex to a boundary vertex, joined by cut arcs, so
//     d_H(s, c) = d_G(s, c) for boundary s, c.  H has ~29 k vertices on the
//     512 x 512 grid cut in 256 parts (G: 262 k), so a source's distance
//     array (118 KB) lives in LDS: one 1024-thread workgroup per source,
//     near-far rounds over the improved parts, whose improved boundary
//     vertices are applied together as rows of L_Q's boundary block (b_Q
//     consecutive targets, read coalesced, one LDS atomic per target).
//  2. expansion.  A path from s to v in part Q enters Q for the last time at
//     a boundary vertex c of Q (or starts there: s is one), then stays in Q:
//     d(s, v) = min over c in B_Q of d_H(s, c) + L_Q(c, v).  Per part that
//     is a min-plus product (sources x b_Q) (x) (b_Q x n_Q) on the product
//     tile, written into the sources' rows at the vertices' own columns.
//
// L_Q for every part is computed once per graph (the index): each part's
// matrix of internal arcs (no boundary rows) closed by the batched closure;
// its boundary rows are kept, boundary columns for H in st2ed order and all
// columns in ascending vertex id for the expansion, so that a tile's 64
// columns fall on few output cache lines.  The results are the reference's
// distances exactly (integer, unique).
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "batch_desc.h"
#include "common.h"
#include "minplus_tile.h"

namespace fapi {

int launch_gather(const fap_graph_dev *g, int32_t start, int32_t n_sub, int32_t n_bdy,
                  const int32_t *rowblock, int64_t ld_rb, int32_t *sub, int64_t ld_sub,
                  hipStream_t s);

}  // namespace fapi

// One part of the index.  H ids of part q's boundary vertices are
// hoff .. hoff + b - 1 (hoff a multiple of 4: a few unused ids pad the gaps,
// so the expansion's A loads are 16-byte aligned).
struct BidxPart {
  int32_t start;    // offset of the part in st2ed
  int32_t n;        // vertices
  int32_t b;        // boundary vertices (the part's first b in st2ed)
  int32_t hoff;     // first H id
  int64_t lbb_off;  // b x b local boundary distances, stride b (H order)
  int64_t lbv_off;  // b x ldv local distances to every vertex, columns by ascending id
  int32_t ldv;      // n rounded up to a multiple of 4
  int32_t pad;
};

struct fap_bidx {
  int32_t K = 0;
  int32_t nbp = 0;         // H ids, padding included
  int32_t nb = 0;          // boundary vertices
  int32_t delta = 0;       // near-far band of the boundary SSSP
  int32_t nctiles = 0;     // expansion column tiles (64 columns of one part)
  std::vector<BidxPart> h_parts;   // host copy, index = part id - 1
  BidxPart *parts = nullptr;       // device, K
  int32_t *hpart = nullptr;        // nbp: part index of an H id (-1 for padding)
  int4 *prec = nullptr;            // K: {hoff, b | x << 11, lbb offset, ext offset}
  int32_t *ext = nullptr;          // per part: x targets (H id | part << 16), then b x x weights
  int32_t *lbb = nullptr;
  int32_t *lbv = nullptr;
  int32_t *colid = nullptr;        // V: vertex ids of each part ascending, at the part's offset
  int2 *ctiles = nullptr;          // {part index, first local column}
  unsigned int *err = nullptr;     // set when a boundary SSSP exceeds its round bound
  size_t bytes = 0;
};

namespace fapk {

constexpr int HS_THREADS = 1024;
constexpr int HS_WAVES = HS_THREADS / 64;
constexpr int HS_MAX_NB = 32768;  // LDS: 128 KB of distances + 4 KB of improved-vertex bits
constexpr int HS_MAX_K = 1024;    // LDS: 16 KB of part records + 128 B of improved-part bits
constexpr int HS_MAX_B = 1024;    // boundary vertices of one part (b < 2^11)
constexpr int HS_MAX_X = 4096;    // targets of one part's cut arcs (x < 2^13)
constexpr int HS_ROWS = 8;        // claimed rows applied together
constexpr int HS_CH = 4;          // 64-column chunks per pass: 32 row loads in flight per lane

// Boundary SSSP, one source per workgroup pass, its distances in LDS, in
// near-far rounds over parts.  Part q's row for boundary vertex c holds the
// distances inside q from c to q's b boundary vertices, then the weights of
// c's cut arcs to the x boundary vertices of other parts that q's cut arcs
// reach (INF where c has no such arc).  Every improved vertex has its bit
// set, and then its part's bit.  A round: the part bits are collected into
// a list (and cleared); the waves take its parts one at a time; for one, a
// wave claims -- a lane per vertex -- the
// part's improved vertices whose distance is below the threshold T (their
// bits cleared; vertices left improved set the part's bit again) and
// applies them together: lane j computes min over claimed c of d[c] +
// row_c[j] for columns j, 8 rows x 4 chunks of 64 columns loaded at a time,
// then lowers each target's distance by one LDS atomic.  Distances are read
// after the claim, so a later decrease sets the bit anew and the vertex is
// applied again.  A round that claims nothing raises T to the smallest
// improved distance + delta (exact: nothing changed during it); a round
// with no improved vertex ends the source.  Every vertex left unimproved
// has had its row applied with its final distance, so the result is exact;
// every distance is the length of a walk, so none is too small.  Rounds are
// bounded (max_rounds: *err is set and every workgroup stops), so the
// launch always drains.
__global__ __launch_bounds__(HS_THREADS) void bidx_sssp_kernel(
    const int32_t *__restrict__ srcs, int32_t nsrc, int32_t nbp, int32_t K,
    const int4 *__restrict__ prec, const int32_t *__restrict__ hpart,
    const int32_t *__restrict__ lbb, const int32_t *__restrict__ ext, int32_t delta,
    int32_t *__restrict__ X, int64_t ldx, int32_t max_rounds, unsigned int *err) {
  __shared__ int32_t dist[HS_MAX_NB];
  __shared__ uint32_t bits[HS_MAX_NB / 32];
  __shared__ int4 pinfo[HS_MAX_K];  // {hoff, b | x << 11, lbb offset, ext offset}
  __shared__ uint32_t pbits[HS_MAX_K / 32];
  __shared__ uint64_t cmask[HS_WAVES][HS_MAX_B / 64];  // a wave's claimed vertices
  __shared__ int32_t plist[HS_MAX_K];                  // the round's improved parts
  __shared__ int32_t s_processed, s_nextmin, s_stop, s_np, s_take;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = (nbp + 31) >> 5, nkw = (K + 31) >> 5;
  for (int i = tid; i < K; i += HS_THREADS) pinfo[i] = prec[i];
  for (int si = blockIdx.x; si < nsrc; si += gridDim.x) {
    if (tid == 0) s_stop = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    for (int i = tid; i < nbp; i += HS_THREADS) dist[i] = FAP_INF;
    for (int i = tid; i < nw; i += HS_THREADS) bits[i] = 0u;
    for (int i = tid; i < nkw; i += HS_THREADS) pbits[i] = 0u;
    __syncthreads();
    if (s_stop) return;  // uniform: another workgroup hit the round bound
    if (tid == 0) {
      const int h0 = srcs[si], q0 = hpart[h0];
      dist[h0] = 0;
      bits[h0 >> 5] = 1u << (h0 & 31);
      pbits[q0 >> 5] = 1u << (q0 & 31);
    }
    int T = delta;
    int rounds = 0;
    bool ok = true;
    for (;;) {
      // the round's list of improved parts (their bits cleared), taken by
      // the waves one part at a time
      if (wave == 0) {
        int base = 0;
        for (int w0 = 0; w0 < nkw; w0 += 64) {
          const int w = w0 + lane;
          uint32_t word = w < nkw ? pbits[w] : 0u;
          if (w < nkw) pbits[w] = 0u;
          const int cnt = __popc(word);
          int pre = cnt;  // inclusive prefix over the lanes
#pragma unroll
          for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(pre, d);
            if (lane >= d) pre += y;
          }
          int slot = base + pre - cnt;
          while (word) {
            plist[slot++] = (w << 5) + __ffs(word) - 1;
            word &= word - 1u;
          }
          base += __shfl(pre, 63);
        }
        if (lane == 0) {
          s_np = base;
          s_take = 0;
          s_processed = 0;
          s_nextmin = FAP_INF;
        }
      }
      __syncthreads();
      int processed = 0, nextmin = FAP_INF;
      const int np = s_np;
      for (;;) {
        int pick = 0;
        if (lane == 0) pick = atomicAdd(&s_take, 1);
        pick = __shfl(pick, 0);
        if (pick >= np) break;
        const int q = plist[pick];
        const int4 pi = pinfo[q];
        const int hoff = pi.x, b = pi.y & 0x7FF, x = pi.y >> 11;
        const int lbb_off = pi.z, ext_off = pi.w;
        // claim, a lane per boundary vertex
        const int nmask = (b + 63) >> 6;
        int nclaimed = 0;
        bool anyleft = false;
        for (int m0 = 0; m0 < nmask; m0++) {
          const int i = (m0 << 6) + lane, v = hoff + i;
          bool cl = false, lf = false;
          if (i < b && ((__hip_atomic_load(&bits[v >> 5], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP) >> (v & 31)) & 1u)) {
            const int dv = dist[v];
            cl = dv < T;
            lf = !cl;
            if (lf) nextmin = min(nextmin, dv);
          }
          const uint64_t m = __ballot(cl);
          if (lane == 0) cmask[wave][m0] = m;
          if (cl) atomicAnd(&bits[v >> 5], ~(1u << (v & 31)));
          nclaimed += __popcll((unsigned long long)m);
          anyleft = anyleft || __ballot(lf) != 0;
        }
        if (anyleft && lane == 0) atomicOr(&pbits[q >> 5], 1u << (q & 31));
        if (nclaimed == 0) continue;
        processed++;
        bool improved = false;
        const int ncol = b + x;
        for (int p0 = 0; p0 < ncol; p0 += 64 * HS_CH) {
          int acc[HS_CH];
#pragma unroll
          for (int ch = 0; ch < HS_CH; ch++) acc[ch] = FAP_INF;
          int mi = 0;
          uint64_t mcur = cmask[wave][0];
          for (;;) {
            int cs[HS_ROWS];
            int k = 0;
#pragma unroll
            for (int g = 0; g < HS_ROWS; g++) {
              while (mcur == 0 && mi + 1 < nmask) mcur = cmask[wave][++mi];
              if (mcur) {
                cs[g] = (mi << 6) + __ffsll((unsigned long long)mcur) - 1;  // row in the part
                mcur &= mcur - 1;
                k = g + 1;
              } else {
                cs[g] = 0;
              }
            }
            if (k == 0) break;
            int w[HS_ROWS][HS_CH];
#pragma unroll
            for (int g = 0; g < HS_ROWS; g++)
#pragma unroll
              for (int ch = 0; ch < HS_CH; ch++) {
                const int col = p0 + ch * 64 + lane;
                w[g][ch] = FAP_INF;
                if (g < k && col < ncol)
                  w[g][ch] = col < b ? lbb[lbb_off + cs[g] * b + col]
                                     : ext[ext_off + x + cs[g] * x + (col - b)];
              }
#pragma unroll
            for (int g = 0; g < HS_ROWS; g++)
              if (g < k) {
                const int dg = dist[hoff + cs[g]];
#pragma unroll
                for (int ch = 0; ch < HS_CH; ch++) acc[ch] = min(acc[ch], dg + w[g][ch]);
              }
            if (k < HS_ROWS) break;
          }
#pragma unroll
          for (int ch = 0; ch < HS_CH; ch++) {
            const int col = p0 + ch * 64 + lane;
            if (col >= ncol) continue;
            int u, q2;
            if (col < b) {
              u = hoff + col;
              q2 = q;
            } else {
              const int packed = ext[ext_off + col - b];
              u = packed & 0xFFFF;
              q2 = packed >> 16;
            }
            if (acc[ch] < dist[u]) {
              const int old = atomicMin(&dist[u], acc[ch]);
              if (acc[ch] < old) {
                atomicOr(&bits[u >> 5], 1u << (u & 31));
                if (q2 == q) improved = true;
                else atomicOr(&pbits[q2 >> 5], 1u << (q2 & 31));
              }
            }
          }
        }
        if (__ballot(improved) && lane == 0) atomicOr(&pbits[q >> 5], 1u << (q & 31));
      }
      if (lane == 0 && processed) atomicAdd(&s_processed, processed);
      if (nextmin < FAP_INF) atomicMin(&s_nextmin, nextmin);  // per lane: lanes read vertices
      __syncthreads();
      const int proc = s_processed, nm = s_nextmin;
      __syncthreads();  // everyone has read both before thread 0 resets them
      if (proc == 0) {
        if (nm >= FAP_INF) break;
        T = nm + delta;
      }
      if (++rounds > max_rounds) {
        ok = false;
        break;
      }
    }
    if (!ok) {
      if (tid == 0) atomicOr(err, 1u);
      return;  // uniform
    }
    int32_t *xr = X + (int64_t)si * ldx;
    for (int i = tid; i < nbp; i += HS_THREADS) xr[i] = dist[i];
    __syncthreads();  // the next source re-initialises the arrays
  }
}

// Rows of the expansion: tile (64 sources, 64 columns of part q) =
// min over c < b_q of X[s][hoff_q + c] + LBV_q[c][j], written to
// rowptr[s][colid[start_q + j]].  Out-of-range sources, columns and c read
// INF; a part without boundary vertices gives INF (it is cut off).
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void bidx_expand_kernel(
    const int32_t *__restrict__ X, int64_t ldx, int32_t nsrc, int32_t *const *__restrict__ rowptr,
    const int2 *__restrict__ ctiles, const BidxPart *__restrict__ parts,
    const int32_t *__restrict__ lbv, const int32_t *__restrict__ colid) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MpSmall::LDS_WORDS];
  const int2 ct = ctiles[blockIdx.x];
  const BidxPart P = parts[ct.x];
  const int j0 = ct.y;
  const int r0 = blockIdx.y * 64;
  const int m_valid = min(64, nsrc - r0), n_valid = min(64, P.n - j0), K = P.b;
  const int32_t *A = X + (int64_t)r0 * ldx + P.hoff;
  const int32_t *B = lbv + P.lbv_off + j0;
  const int flags = MP_VEC_A | MP_VEC_B;  // hoff, ldx, lbv_off, ldv, j0 are multiples of 4
  int acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[i][c] = kInf;
  int ra[4], rb[4];
  const int nch = (K + MP_BK - 1) / MP_BK;
  int32_t *As = lds;
  int32_t *Bs = lds + 64 * MP_AS_LD;
  if (nch > 0) {
    MpSmall::load_chunk(A, ldx, B, P.ldv, m_valid, n_valid, K, 0, flags, ra, rb);
    MpSmall::store_chunk(As, Bs, ra, rb);
  }
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const bool more = ch + 1 < nch;
    if (more) MpSmall::load_chunk(A, ldx, B, P.ldv, m_valid, n_valid, K, (ch + 1) * MP_BK, flags,
                                  ra, rb);
    const int krem = K - ch * MP_BK;
    MpSmall::compute(As, Bs, krem >= MP_BK ? MP_BK : (krem + 1) & ~1, acc);
    __syncthreads();
    if (more) {
      MpSmall::store_chunk(As, Bs, ra, rb);
      __syncthreads();
    }
  }
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int row0 = (wave >> 1) * 32 + (lane >> 3);
  const int col0 = (wave & 1) * 32 + (lane & 7) * 4;
  const int32_t *ids = colid + P.start + j0;
  int cid[4];
#pragma unroll
  for (int c = 0; c < 4; c++) cid[c] = col0 + c < n_valid ? ids[col0 + c] : 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int row = row0 + 8 * i;
    if (row >= m_valid) continue;
    int32_t *out = rowptr[r0 + row];
#pragma unroll
    for (int c = 0; c < 4; c++)
      if (col0 + c < n_valid) out[cid[c]] = acc[i][c];
  }
}

// Index extraction for one part: from its closed local matrix (st2ed
// order, n x n, stride ld) keep the boundary rows -- columns < b as the
// part's block of H arcs, every column reordered by ascending vertex id for
// the expansion.
__global__ void bidx_extract_kernel(const int32_t *__restrict__ sub, int64_t ld, BidxPart P,
                                    const int32_t *__restrict__ colid,
                                    const int32_t *__restrict__ ed2st, int32_t *__restrict__ lbb,
                                    int32_t *__restrict__ lbv) {
  const int i = blockIdx.y;  // boundary row
  const int32_t *srow = sub + (int64_t)i * ld;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < P.ldv; j += gridDim.x * blockDim.x) {
    if (j < P.b) lbb[P.lbb_off + (int64_t)i * P.b + j] = srow[j];
    lbv[P.lbv_off + (int64_t)i * P.ldv + j] =
        j < P.n ? srow[ed2st[colid[P.start + j]] - P.start] : kInf;
  }
}

}  // namespace fapk

namespace fapi {

namespace {

void free_bidx(fap_bidx *x) {
  if (!x) return;
  void *bufs[] = {x->parts, x->hpart, x->prec,  x->ext,    x->lbb,
                  x->lbv,   x->colid, x->ctiles, x->err};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete x;
}

template <typename T>
int upload(const std::vector<T> &h, T **d, size_t *bytes, hipStream_t s) {
  const size_t nb = std::max<size_t>(h.size(), 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)d, nb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *d = nullptr;
    return fail(FAP_E_NOMEM, "boundary index: device allocation of " + std::to_string(nb) +
                                 " bytes failed");
  }
  *bytes += nb;
  if (!h.empty()) FAP_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T),
                                         hipMemcpyHostToDevice, s));
  return FAP_OK;
}

}  // namespace

void bidx_release(fap_graph_dev *g) {
  free_bidx(g->bidx);
  g->bidx = nullptr;
  g->bidx_state = 0;
}

// Builds the index of g (partition attached).  Leaves g->bidx null (and
// returns FAP_OK) when the graph does not qualify: more boundary vertices
// than the LDS distance array holds, boundary vertices not first in their
// part's st2ed range, or duplicate arcs (the part matrix keeps the last
// duplicate, as the reference's does, which need not be the shortest).
// Synchronous.
int bidx_build(fap_ctx *ctx, fap_graph_dev *g) {
  if (g->bidx_state) return FAP_OK;
  g->bidx_state = -1;
  const int32_t V = g->V;
  if (V <= 0 || (int64_t)g->h_st2ed.size() != V || (int64_t)g->h_graph_id.size() != V)
    return FAP_OK;
  hipStream_t s = ctx->stream;
  // the CSR back on the host
  std::vector<int32_t> ro(V + 1);
  std::vector<int2> ed((size_t)g->E);
  FAP_HIP(hipMemcpyAsync(ro.data(), g->row_offset, sizeof(int32_t) * (V + 1),
                         hipMemcpyDeviceToHost, s));
  if (g->E) FAP_HIP(hipMemcpyAsync(ed.data(), g->edges, sizeof(int2) * g->E, hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  const std::vector<int32_t> &st2ed = g->h_st2ed, &gid = g->h_graph_id;
  int32_t K = 0;
  for (int32_t v = 0; v < V; v++) K = std::max(K, gid[v]);
  if (K <= 0) return FAP_OK;
  // boundary flags (both ends of every cut arc) and duplicate arcs
  std::vector<uint8_t> bdy(V, 0);
  std::vector<int32_t> seen(V, -1);
  for (int32_t u = 0; u < V; u++) {
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      const int32_t v = ed[e].x;
      if (v < 0 || v >= V) return FAP_OK;
      if (seen[v] == u) return FAP_OK;  // duplicate arc
      seen[v] = u;
      if (gid[v] != gid[u]) bdy[u] = bdy[v] = 1;
    }
  }
  // parts: contiguous ranges of st2ed, boundary first
  fap_bidx *x = new fap_bidx();
  x->K = K;
  x->h_parts.assign(K, BidxPart{});
  std::vector<int32_t> cnt(K + 1, 0), bcnt(K + 1, 0);
  for (int32_t v = 0; v < V; v++) {
    if (gid[v] < 1 || gid[v] > K) {
      delete x;
      return FAP_OK;
    }
    cnt[gid[v]]++;
    bcnt[gid[v]] += bdy[v];
  }
  int64_t hoff = 0, lbb = 0, lbv = 0;
  int32_t pos = 0;
  for (int32_t q = 1; q <= K; q++) {
    BidxPart &P = x->h_parts[q - 1];
    P.start = pos;
    P.n = cnt[q];
    P.b = bcnt[q];
    for (int32_t i = 0; i < P.n; i++) {
      const int32_t v = st2ed[pos + i];
      if (gid[v] != q || (i < P.b) != (bdy[v] != 0)) {
        delete x;
        return FAP_OK;
      }
    }
    P.hoff = (int32_t)hoff;
    hoff += (P.b + 3) & ~3;
    P.ldv = (P.n + 3) & ~3;
    P.lbb_off = lbb;
    lbb += (int64_t)P.b * P.b;
    P.lbv_off = lbv;
    lbv += (int64_t)P.b * P.ldv;
    pos += P.n;
  }
  x->nbp = (int32_t)hoff;
  int32_t bmax = 0;
  for (const BidxPart &P : x->h_parts) bmax = std::max(bmax, P.b);
  if (x->nbp > fapk::HS_MAX_NB || x->nbp == 0 || K > fapk::HS_MAX_K || bmax > fapk::HS_MAX_B) {
    delete x;
    return FAP_OK;
  }
  x->nb = 0;
  for (int32_t q = 1; q <= K; q++) x->nb += bcnt[q];
  // H ids, cut arcs, columns by ascending id, expansion tiles
  std::vector<int32_t> hid(V, -1), hpart(x->nbp, -1);
  for (int32_t q = 0; q < K; q++) {
    const BidxPart &P = x->h_parts[q];
    for (int32_t i = 0; i < P.b; i++) {
      hid[st2ed[P.start + i]] = P.hoff + i;
      hpart[P.hoff + i] = q;
    }
  }
  // each part's cut arcs as extra columns of its rows: the distinct targets
  // (H id | part << 16), then a b x x block of weights (INF: no arc)
  std::vector<int32_t> ext;
  std::vector<int4> prec(K);
  for (int32_t q = 0; q < K; q++) {
    const BidxPart &P = x->h_parts[q];
    std::vector<int32_t> tg;  // H ids of the targets
    for (int32_t i = 0; i < P.b; i++) {
      const int32_t u = st2ed[P.start + i];
      for (int32_t e = ro[u]; e < ro[u + 1]; e++)
        if (gid[ed[e].x] != gid[u]) tg.push_back(hid[ed[e].x]);
    }
    std::sort(tg.begin(), tg.end());
    tg.erase(std::unique(tg.begin(), tg.end()), tg.end());
    const int32_t nx = (int32_t)tg.size();
    if (nx > fapk::HS_MAX_X || (int64_t)ext.size() + nx + (int64_t)P.b * nx > INT32_MAX ||
        P.lbb_off > INT32_MAX) {
      delete x;
      return FAP_OK;
    }
    prec[q] = make_int4(P.hoff, P.b | (nx << 11), (int32_t)P.lbb_off, (int32_t)ext.size());
    for (int32_t t : tg) ext.push_back(t | (hpart[t] << 16));
    const size_t wbase = ext.size();
    ext.resize(wbase + (size_t)P.b * nx, FAP_INF);
    for (int32_t i = 0; i < P.b; i++) {
      const int32_t u = st2ed[P.start + i];
      for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
        const int32_t v = ed[e].x;
        if (gid[v] == gid[u]) continue;
        const int32_t xi = (int32_t)(std::lower_bound(tg.begin(), tg.end(), hid[v]) - tg.begin());
        ext[wbase + (size_t)i * nx + xi] = ed[e].y;
      }
    }
  }
  std::vector<int32_t> colid(st2ed);
  std::vector<int2> ctiles;
  int64_t wsum = 0;
  for (int64_t e = 0; e < g->E; e++) wsum += ed[e].y;
  for (int32_t q = 0; q < K; q++) {
    const BidxPart &P = x->h_parts[q];
    std::sort(colid.begin() + P.start, colid.begin() + P.start + P.n);
    for (int32_t j = 0; j < P.n; j += 64) ctiles.push_back(make_int2(q, j));
  }
  x->nctiles = (int32_t)ctiles.size();
  // near-far band: 64 mean arc weights -- a round applies whole parts, so a
  // wide band (fewer rounds, more vertices applied again) pays
  const double mean_w = g->E ? (double)wsum / (double)g->E : 1.0;
  x->delta = (int32_t)std::min<double>(std::max(1.0, 64.0 * mean_w), (double)(FAP_INF / 4));
  int rc = FAP_OK;
  if (rc == FAP_OK) rc = upload(x->h_parts, &x->parts, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(hpart, &x->hpart, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(prec, &x->prec, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(ext, &x->ext, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(colid, &x->colid, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(ctiles, &x->ctiles, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(std::vector<unsigned int>(1, 0u), &x->err, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(std::vector<int32_t>((size_t)std::max<int64_t>(lbb, 1)), &x->lbb,
                                &x->bytes, s);
  if (rc == FAP_OK) rc = upload(std::vector<int32_t>((size_t)std::max<int64_t>(lbv, 1)), &x->lbv,
                                &x->bytes, s);
  // local closures, a few parts at a time (<= 256 MiB of part matrices)
  for (int32_t q0 = 0; q0 < K && rc == FAP_OK;) {
    int32_t q1 = q0;
    int64_t words = 0;
    std::vector<int64_t> off;
    while (q1 < K) {
      const int64_t ld = (x->h_parts[q1].n + 3) & ~3;
      const int64_t w = ld * x->h_parts[q1].n;
      if (q1 > q0 && (words + w) * 4 > ((int64_t)256 << 20)) break;
      off.push_back(words);
      words += w;
      q1++;
    }
    rc = ensure_ws(ctx, ctx->ws_sub, (size_t)std::max<int64_t>(words, 1) * sizeof(int32_t));
    if (rc != FAP_OK) break;
    const int32_t cnt_p = q1 - q0;
    std::vector<int32_t *> subs(cnt_p);
    std::vector<int32_t> ns(cnt_p);
    std::vector<int64_t> lds(cnt_p);
    for (int32_t i = 0; i < cnt_p; i++) {
      const BidxPart &P = x->h_parts[q0 + i];
      subs[i] = (int32_t *)ctx->ws_sub.ptr + off[i];
      ns[i] = P.n;
      lds[i] = (P.n + 3) & ~3;
      rc = launch_gather(g, P.start, P.n, 0, nullptr, 0, subs[i], lds[i], s);
      if (rc != FAP_OK) break;
    }
    if (rc == FAP_OK) rc = fap_fw_batched(ctx, cnt_p, subs.data(), ns.data(), lds.data(), s);
    for (int32_t i = 0; i < cnt_p && rc == FAP_OK; i++) {
      const BidxPart &P = x->h_parts[q0 + i];
      if (P.b == 0) continue;
      hipLaunchKernelGGL(fapk::bidx_extract_kernel, dim3((P.ldv + 255) / 256, P.b), dim3(256), 0,
                         s, subs[i], lds[i], P, x->colid, g->ed2st, x->lbb, x->lbv);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = hip_fail(e, "boundary index extract", __FILE__, __LINE__);
    }
    if (rc == FAP_OK) {
      const hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(e, "boundary index build", __FILE__, __LINE__);
    }
    q0 = q1;
  }
  if (rc != FAP_OK) {
    (void)hipStreamSynchronize(s);
    free_bidx(x);
    return rc;
  }
  g->bidx = x;
  g->bidx_state = 1;
  return FAP_OK;
}

// Stage-1 rows of `nsrc` boundary sources through the index: d_hsrc holds
// their H ids (device), d_rowptr their output rows (device row pointers,
// V columns each).  Uses ctx->ws_bidx (nsrc x nbp int32).  Asynchronous on s.
int bidx_rows(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_hsrc, int32_t nsrc,
              int32_t *const *d_rowptr, hipStream_t s) {
  const fap_bidx *x = g->bidx;
  if (!x) return fail(FAP_E_STATE, "no boundary index");
  if (nsrc <= 0) return FAP_OK;
  const int64_t ldx = x->nbp;  // a multiple of 4
  FAP_TRY(ensure_ws(ctx, ctx->ws_bidx, (size_t)nsrc * ldx * sizeof(int32_t)));
  int32_t *X = (int32_t *)ctx->ws_bidx.ptr;
  const int grid = std::min<int>(nsrc, ctx->num_cus);
  const int32_t max_rounds = 64 * x->nbp + 4096;
  hipLaunchKernelGGL(fapk::bidx_sssp_kernel, dim3(grid), dim3(fapk::HS_THREADS), 0, s, d_hsrc,
                     nsrc, x->nbp, x->K, x->prec, x->hpart, x->lbb, x->ext,
                     ctx->opt_bidx_delta > 0 ? ctx->opt_bidx_delta : x->delta, X, ldx, max_rounds,
                     x->err);
  FAP_CHECK_LAUNCH();
  const int rows = (nsrc + 63) / 64;
  if (rows > 65535) return fail(FAP_E_ARG, "boundary rows: too many sources in one call");
  hipLaunchKernelGGL(fapk::bidx_expand_kernel, dim3(x->nctiles, rows), dim3(fapk::MP_THREADS), 0,
                     s, X, ldx, nsrc, d_rowptr, x->ctiles, x->parts, x->lbv, x->colid);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int32_t bidx_source_base(const fap_graph_dev *g, int32_t start, int32_t n, int32_t b) {
  const fap_bidx *x = g->bidx;
  if (!x || start < 0 || start >= g->V || n <= 0) return -1;
  const int32_t q = g->h_graph_id[g->h_st2ed[start]] - 1;
  if (q < 0 || q >= x->K) return -1;
  const BidxPart &P = x->h_parts[q];
  return (P.start == start && P.n == n && P.b == b) ? P.hoff : -1;
}

// Raised bound flag of the boundary SSSP (a round count no exact run
// reaches; reported as an error, the rows are not valid then).
int bidx_check(const fap_graph_dev *g) {
  if (!g->bidx) return FAP_OK;
  unsigned int h = 0;
  FAP_HIP(hipMemcpy(&h, g->bidx->err, sizeof(h), hipMemcpyDeviceToHost));
  if (h) return fail(FAP_E_DEVICE, "boundary SSSP exceeded its round bound");
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_graph_build_index(fap_ctx *ctx, fap_graph_dev *g, int32_t *built) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  FAP_HIP(hipSetDevice(ctx->device));
  FAP_TRY(bidx_build(ctx, g));
  if (built) *built = g->bidx ? 1 : 0;
  return FAP_OK;
}

extern "C" int fap_graph_index_info(const fap_graph_dev *g, int32_t *boundary, int64_t *bytes) {
  if (!g) return fail(FAP_E_ARG, "null graph");
  if (boundary) *boundary = g->bidx ? g->bidx->nb : 0;
  if (bytes) *bytes = g->bidx ? (int64_t)g->bidx->bytes : 0;
  return FAP_OK;
}
