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
//     d_H(s, c) = d_G(s, c) for boundary s, c.  All boundary vertices of a
//     part P are solved together, as a block of rows X_P (b_P x all boundary
//     vertices), by Bellman-Ford over the parts: starting from X_P[., B_P] =
//     L_P on P's own boundary, a block X_P[., B_Q] is updated after a block
//     X_P[., B_Q'] with a cut arc Q' -> Q changed: X_P[., B_Q] =
//     min(X_P[., B_Q], E (x) L_Q[k, .]) over the border positions k of Q
//     entered from such parts Q', where E[., k] is the best way into k over
//     the cut arcs c' -> k (min of X_P[., c'] + w) -- a min-plus product with
//     b_P rows on the product tile, the cut-arc minimum gathered while the A
//     operand is staged.  Changed blocks are propagated in distance bands
//     (delta-stepping over blocks).  When nothing is pending every arc of H
//     is satisfied, so X_P is exact (the values are lengths of walks, never
//     too small).  The iterations run on the device: a one-workgroup
//     scheduling step lists each iteration's tiles, and the host only
//     checks for convergence every few iterations.  On the 512x512 grid cut
//     in 256 parts, 32 parts per call: ~224 iterations.
//  2. expansion.  A path from s to v in part Q enters Q for the last time at
//     a boundary vertex c of Q (or starts there: s is one), then stays in Q:
//     d(s, v) = min over c in B_Q of d_H(s, c) + L_Q(c, v).  Per part that
//     is a min-plus product (sources x b_Q) (x) (b_Q x n_Q) on the product
//     tile, written into the sources' rows at the vertices' own columns.
//
// L_Q for every part is computed once per graph (the index): each part's
// matrix of internal arcs (no boundary rows) closed by the batched closure;
// its boundary rows are kept, in H order: boundary columns for step 1 and
// all columns in ascending vertex id for the expansion, so that a tile's
// columns fall on few output cache lines.  The results are the reference's
// distances exactly (integer, unique).
#include <math.h>
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
// so row blocks of X start 16-byte aligned), in border order: grouped by
// the neighbouring part each vertex has its first cut arc with, so that the
// vertices on either side of one border take consecutive ids.  The
// position of a vertex in that order is its H position (H id - hoff).
struct BidxPart {
  int32_t start;    // offset of the part in st2ed
  int32_t n;        // vertices
  int32_t b;        // boundary vertices (the part's first b in st2ed)
  int32_t hoff;     // first H id
  int64_t lbb_off;  // b x ldb local boundary distances (H order)
  int64_t lbv_off;  // b x ldv local distances to every vertex, columns by ascending id
  int32_t ldb;      // b rounded up to a multiple of 4
  int32_t ldv;      // n rounded up to a multiple of 4
};

// One tile (64 rows x 64 or 128 columns) of a block update: rows tr*64.. of
// the block whose first row of X is row0 (rows rows), columns tc*64.. or
// tc*128.. of part q's boundary; the
// product runs over the border positions of q entered from the parts
// whose bits are set in `slots` (bit s: q's s-th in-neighbour part; bit 31
// stands for every in-neighbour from the 31st on).  `blk` is the block's
// index p * K + q in the pending array.
struct BidxTile {
  int32_t row0, rows, q, tr, tc, blk;
  uint32_t slots;
  int32_t pad;
};

// Control words of the block iterations (device).
enum { BC_THR = 0, BC_NTILES = 2, BC_DONE = 3, BC_WORDS = 16 };

// Longest border-position list a tile may take (the sum of a part's
// in-border lengths); a graph with a longer one gets no index.
constexpr int kBidxKlMax = 2048;

// Expansion tiles: 16 kExpandTM sources x kExpandCols columns of a part.
constexpr int kExpandTM = 8, kExpandTN = 8, kExpandCols = 16 * kExpandTN;

struct fap_bidx {
  int32_t K = 0;
  int32_t nbp = 0;         // H ids, padding included
  int32_t nb = 0;          // boundary vertices
  int32_t nctiles = 0;     // expansion column tiles (kExpandCols columns of one part)
  int32_t delta = 1;       // distance band of the block iterations
  std::vector<BidxPart> h_parts;          // host copy, index = part id - 1
  std::vector<int32_t> h_hpos;            // H position of each boundary vertex (local index order)
  std::vector<int32_t> boff;              // K + 1: each part's first entry in h_hpos
  BidxPart *parts = nullptr;       // device, K
  int32_t *in_off = nullptr;       // nbp + 1: cut arcs into each H id
  int2 *in_arc = nullptr;          // {source H id, weight}
  int2 *in1 = nullptr;             // nbp: {source H id, weight} of an H id's only cut arc,
                                   // else {-1 - first arc, end of its arcs} (none: {-1, 0})
  int32_t *hloc = nullptr;         // nbp: local (st2ed) index of each H id's vertex
  // borders: part q's in-neighbour slots inb_off[q] .. inb_off[q + 1] - 1;
  // slot e's H positions of q (entered by cut arcs from that neighbour,
  // ascending) at inb_pos[inb_pos_off[e] .. inb_pos_off[e + 1] - 1]; part
  // q's out-neighbours outb[outb_off[q] ..] = {part q2, q's slot index among
  // q2's in-neighbours}
  int32_t *inb_off = nullptr;      // K + 1
  int32_t *inb_pos_off = nullptr;  // slots + 1
  int32_t *inb_pos = nullptr;
  int32_t *outb_off = nullptr;     // K + 1
  int2 *outb = nullptr;
  int32_t *lbb = nullptr;
  int32_t *lbv = nullptr;
  int32_t *colid = nullptr;        // V: vertex ids of each part ascending, at the part's offset
  int2 *ctiles = nullptr;          // {part index, first local column}
  size_t bytes = 0;
};

namespace fapk {

// Start of a block: row r of X (part q, source at H position i) = INF,
// with L_q(i, .) on q's own boundary columns.
__global__ void bidx_block_init_kernel(int32_t *__restrict__ X, int64_t ldx, int32_t nbp,
                                       const int2 *__restrict__ rowq,
                                       const BidxPart *__restrict__ parts,
                                       const int32_t *__restrict__ lbb) {
  const int r = blockIdx.x;
  const int2 qi = rowq[r];
  const BidxPart P = parts[qi.x];
  int32_t *xr = X + (int64_t)r * ldx;
  for (int c = threadIdx.x; c < nbp; c += blockDim.x) {
    const int j = c - P.hoff;
    xr[c] = (j >= 0 && j < P.b) ? lbb[P.lbb_off + (int64_t)qi.y * P.ldb + j] : kInf;
  }
}

// LDS copies the scheduling step keeps when they fit (global otherwise):
// the marks of up to kSchedBlocks blocks, boundary sizes and out-neighbour
// offsets of up to kSchedParts parts.
constexpr int kSchedBlocks = 8192, kSchedParts = 2048;

// Scheduling step of the block iterations (one workgroup).  pend[p * K + q]
// is a block's smallest changed value (INF: not pending; the block launch
// lowers it as it improves values).  The step raises the band threshold
// past the smallest pending value when none lies below it, takes every
// pending block below the threshold (clearing it), marks for each part q2
// that a cut arc from the block's part q enters the slot of q among q2's
// in-neighbours (mask[p * K + q2]), and lists the 64 x tcol tiles of the
// marked blocks.  With nothing pending it sets the done word and lists
// nothing, so launches enqueued after convergence are no-ops.
__global__ __launch_bounds__(1024) void bidx_schedule_kernel(
    int32_t *ctrl, int32_t *pend, uint32_t *mask, BidxTile *tiles, const int2 *__restrict__ prow, int32_t count, int32_t K, int32_t delta,
    int32_t tcol, const BidxPart *__restrict__ parts, const int32_t *__restrict__ outb_off,
    const int2 *__restrict__ outb) {
  __shared__ int s_red[16];
  __shared__ int s_thr, s_ntl;
  __shared__ uint32_t smask[kSchedBlocks];
  __shared__ int32_t sb[kSchedParts];
  __shared__ int32_t soff[kSchedParts + 1];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nblk = count * K;
  const bool lmask = nblk <= kSchedBlocks, lparts = K <= kSchedParts;
  if (lmask)
    for (int i = tid; i < nblk; i += nthr) smask[i] = 0;
  if (lparts) {
    for (int q = tid; q < K; q += nthr) {
      sb[q] = parts[q].b;
      soff[q] = outb_off[q];
    }
    if (tid == 0) soff[K] = outb_off[K];
  }
  __syncthreads();
  int lo = kInf;
  for (int i = tid; i < nblk; i += nthr) lo = min(lo, pend[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lo = min(lo, __shfl_xor(lo, off));
  if ((tid & 63) == 0) s_red[tid >> 6] = lo;
  __syncthreads();
  if (tid == 0) {
    int m = kInf;
    for (int w = 0; w < (nthr + 63) / 64; w++) m = min(m, s_red[w]);
    int thr = ctrl[BC_THR];
    if (m < kInf && m >= thr) thr = (int)(((int64_t)m / delta + 1) * delta);
    s_thr = m < kInf ? thr : -1;
    ctrl[BC_THR] = thr;
    ctrl[BC_DONE] = m < kInf ? 0 : 1;
    s_ntl = 0;
  }
  __syncthreads();
  const int thr = s_thr;
  if (thr < 0) {
    if (tid == 0) ctrl[BC_NTILES] = 0;
    return;
  }
  for (int i = tid; i < nblk; i += nthr) {
    if (pend[i] >= thr) continue;
    pend[i] = kInf;
    const int p = i / K, q = i - p * K;
    const int e0 = lparts ? soff[q] : outb_off[q], e1 = lparts ? soff[q + 1] : outb_off[q + 1];
    for (int e = e0; e < e1; e++) {
      const int2 o = outb[e];
      const int b = lparts ? sb[o.x] : parts[o.x].b;
      if (b <= 0) continue;
      const uint32_t bit = 1u << min(o.y, 31);
      if (lmask) atomicOr(&smask[p * K + o.x], bit);
      else atomicOr(&mask[p * K + o.x], bit);
    }
  }
  if (!lmask) __threadfence();
  __syncthreads();
  for (int i = tid; i < nblk; i += nthr) {
    uint32_t m;
    if (lmask) {
      m = smask[i];
    } else {
      m = mask[i];
      if (m) mask[i] = 0;
    }
    if (!m) continue;
    const int p = i / K, q2 = i - p * K;
    const int2 pr = prow[p];
    const int rt = (pr.y + 63) >> 6, ct = ((lparts ? sb[q2] : parts[q2].b) + tcol - 1) / tcol;
    const int base = atomicAdd(&s_ntl, rt * ct);
    for (int tr = 0; tr < rt; tr++)
      for (int tc = 0; tc < ct; tc++)
        tiles[base + tr * ct + tc] = BidxTile{pr.x, pr.y, q2, tr, tc, i, m, 0};
  }
  __syncthreads();
  if (tid == 0) ctrl[BC_NTILES] = s_ntl;
}

// Block updates of one iteration, the tiles listed by the scheduling step
// dealt grid-stride.  Tile: X_P[., B_q] = min(X_P[., B_q], E (x) L_q[k, .])
// over the border positions k of q entered from the marked in-neighbours
// (collected in LDS first); A (64 rows x k) is E, staged chunk by chunk
// straight from the cut arcs into q (E[i][k] = min over arcs c' -> hoff_q +
// k of X[i][c'] + w); B is the listed rows of L_q's boundary block; C is X,
// kept where not improved.  The block's pending word takes the smallest
// improved value.  Positions left out are entered only from parts whose
// blocks did not change since this block was last updated with them.  X
// values read by the gather may be lowered concurrently by other tiles;
// either value is the length of a walk, and a lowered one is pending again,
// so the iteration only converges on exact distances.
template <int TM, int TN>
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(TN == 4 ? 5 : 4))) void bidx_block_kernel(
    const int32_t *__restrict__ ctrl, const BidxTile *__restrict__ tiles, int32_t *X, int64_t ldx,
    const BidxPart *__restrict__ parts, const int2 *__restrict__ in1,
    const int2 *__restrict__ in_arc, const int32_t *__restrict__ lbb,
    const int32_t *__restrict__ inb_off, const int32_t *__restrict__ inb_pos_off,
    const int32_t *__restrict__ inb_pos, int32_t *pend) {
  using Tile = MpTile<TM, TN>;
  constexpr int TR = Tile::BM, TC = Tile::BN;  // rows, columns of a tile
  __shared__ __attribute__((aligned(16))) int32_t lds[Tile::LDS_WORDS];
  __shared__ int32_t klist[kBidxKlMax];
  __shared__ int2 kinf[kBidxKlMax];  // in1 of each listed position
  const int ntiles = ctrl[BC_NTILES];
  const int t = threadIdx.x;
  const int arow = t / Tile::A_TPR, akq = (t % Tile::A_TPR) * TM;  // A staging (the tile's layout)
  const int wave = t >> 6, lane = t & 63;
  const int row0 = (wave >> 1) * (8 * TM) + (lane >> 3);  // the tile's output layout
  const int col0 = (wave & 1) * (8 * TN) + (lane & 7) * 4;
  for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const BidxTile T = tiles[ti];
    const BidxPart Q = parts[T.q];
    // border positions of the marked in-neighbours (a corner vertex may
    // appear twice: min-plus does not mind)
    int kn = 0;
    const int s0 = inb_off[T.q], s1 = inb_off[T.q + 1];
    for (int sl = s0; sl < s1; sl++) {
      const int si = sl - s0;
      if (!((T.slots >> min(si, 31)) & 1u)) continue;
      const int p0 = inb_pos_off[sl], len = inb_pos_off[sl + 1] - p0;
      for (int i = t; i < len; i += MP_THREADS) {
        const int pos = inb_pos[p0 + i];
        klist[kn + i] = pos;
        kinf[kn + i] = in1[Q.hoff + pos];
      }
      kn += len;
    }
    const int K = kn;
    __syncthreads();
    const int r0 = T.tr * TR, c0 = T.tc * TC;
    const int m_valid = min(TR, T.rows - r0), n_valid = min(TC, Q.b - c0);
    int32_t *Xt = X + (int64_t)(T.row0 + r0) * ldx;
    const int32_t *B = lbb + Q.lbb_off + c0;
    const int32_t *xa = Xt + (int64_t)arow * ldx;
    int ra[TM], rb[TN];
    auto load = [&](int k0) {
#pragma unroll
      for (int j = 0; j < TM; j++) {
        const int k = k0 + akq + j;
        int v = kInf;
        if (arow < m_valid && k < K) {
          const int2 f = kinf[k];
          if (f.x >= 0) {
            v = min(kInf, xa[f.x] + f.y);  // the position's only cut arc (the common case)
          } else {
            for (int e = -1 - f.x; e < f.y; e++) {
              const int2 a = in_arc[e];
              v = min(v, xa[a.x] + a.y);
            }
          }
        }
        ra[j] = v;
      }
      const int kr = k0 + (t >> 4), col = (t & 15) * TN;
      const int32_t *pb = B + (int64_t)(kr < K ? klist[kr] : 0) * Q.ldb + col;
      if (kr < K && col + TN <= n_valid) {
#pragma unroll
        for (int q = 0; q < TN; q += 4) {
          const int4 v = *reinterpret_cast<const int4 *>(pb + q);
          rb[q] = v.x; rb[q + 1] = v.y; rb[q + 2] = v.z; rb[q + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int j = 0; j < TN; j++) rb[j] = (kr < K && col + j < n_valid) ? pb[j] : kInf;
      }
    };
    int acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int c = 0; c < TN; c++) acc[i][c] = kInf;
    int32_t *As = lds;
    int32_t *Bs = lds + TR * MP_AS_LD;
    const int nch = (K + MP_BK - 1) / MP_BK;
    if (nch > 0) {
      load(0);
      Tile::store_chunk(As, Bs, ra, rb);
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
      const bool more = ch + 1 < nch;
      if (more) load((ch + 1) * MP_BK);
      const int krem = K - ch * MP_BK;
      Tile::compute(As, Bs, krem >= MP_BK ? MP_BK : (krem + 1) & ~1, acc);
      __syncthreads();
      if (more) {
        Tile::store_chunk(As, Bs, ra, rb);
        __syncthreads();
      }
    }
    int lowest = kInf;  // smallest improved value
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const int row = row0 + 8 * i;
      if (row >= m_valid) continue;
      int32_t *cr = Xt + (int64_t)row * ldx + Q.hoff + c0;
#pragma unroll
      for (int h = 0; h < TN / 4; h++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int col = col0 + 32 * h + c;
          if (col >= n_valid) continue;
          if (acc[i][4 * h + c] < cr[col]) {
            cr[col] = acc[i][4 * h + c];
            lowest = min(lowest, acc[i][4 * h + c]);
          }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lowest = min(lowest, __shfl_xor(lowest, off));
    if (lane == 0 && lowest < kInf) atomicMin(&pend[T.blk], lowest);
    __syncthreads();  // klist and the staging are rewritten by the next tile
  }
}

// Rows of the expansion: tile (16 TM sources, 16 TN columns of part q) =
// min over c < b_q of X[s][hoff_q + c] + LBV_q[c][j], written to
// rowptr[s][colid[start_q + j]].  Out-of-range sources, columns and c read
// INF; a part without boundary vertices gives INF (it is cut off).
template <int TM, int TN>
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(TN == 4 ? 8 : 4))) void bidx_expand_kernel(
    const int32_t *__restrict__ X, int64_t ldx, int32_t nsrc, int32_t *const *__restrict__ rowptr,
    const int2 *__restrict__ ctiles, const BidxPart *__restrict__ parts,
    const int32_t *__restrict__ lbv, const int32_t *__restrict__ colid) {
  using Tile = MpTile<TM, TN>;
  constexpr int TR = Tile::BM, TC = Tile::BN;
  __shared__ __attribute__((aligned(16))) int32_t lds[Tile::LDS_WORDS];
  const int2 ct = ctiles[blockIdx.x];
  const BidxPart P = parts[ct.x];
  const int j0 = ct.y;
  const int r0 = blockIdx.y * TR;
  const int m_valid = min(TR, nsrc - r0), n_valid = min(TC, P.n - j0), K = P.b;
  const int32_t *A = X + (int64_t)r0 * ldx + P.hoff;
  const int32_t *B = lbv + P.lbv_off + j0;
  const int flags = MP_VEC_A | MP_VEC_B;  // hoff, ldx, lbv_off, ldv, j0 are multiples of 4
  int acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int c = 0; c < TN; c++) acc[i][c] = kInf;
  int ra[TM], rb[TN];
  const int nch = (K + MP_BK - 1) / MP_BK;
  int32_t *As = lds;
  int32_t *Bs = lds + TR * MP_AS_LD;
  if (nch > 0) {
    Tile::load_chunk(A, ldx, B, P.ldv, m_valid, n_valid, K, 0, flags, ra, rb);
    Tile::store_chunk(As, Bs, ra, rb);
  }
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const bool more = ch + 1 < nch;
    if (more) Tile::load_chunk(A, ldx, B, P.ldv, m_valid, n_valid, K, (ch + 1) * MP_BK, flags,
                               ra, rb);
    const int krem = K - ch * MP_BK;
    Tile::compute(As, Bs, krem >= MP_BK ? MP_BK : (krem + 1) & ~1, acc);
    __syncthreads();
    if (more) {
      Tile::store_chunk(As, Bs, ra, rb);
      __syncthreads();
    }
  }
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int row0 = (wave >> 1) * (8 * TM) + (lane >> 3);  // the tile's output layout
  const int col0 = (wave & 1) * (8 * TN) + (lane & 7) * 4;
  const int32_t *ids = colid + P.start + j0;
#pragma unroll
  for (int h = 0; h < TN / 4; h++) {
    int cid[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int col = col0 + 32 * h + c;
      cid[c] = col < n_valid ? ids[col] : -1;
    }
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const int row = row0 + 8 * i;
      if (row >= m_valid) continue;
      int32_t *out = rowptr[r0 + row];
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (cid[c] >= 0) out[cid[c]] = acc[i][4 * h + c];
    }
  }
}

// Index extraction for one part: from its closed local matrix (st2ed
// order, n x n, stride ld) keep the boundary rows, in H order: columns of
// boundary vertices in H order as the part's boundary block (stride ldb),
// every column by ascending vertex id for the expansion.
__global__ void bidx_extract_kernel(const int32_t *__restrict__ sub, int64_t ld, BidxPart P,
                                    const int32_t *__restrict__ hloc,
                                    const int32_t *__restrict__ colid,
                                    const int32_t *__restrict__ ed2st, int32_t *__restrict__ lbb,
                                    int32_t *__restrict__ lbv) {
  const int i = blockIdx.y;  // H position of the row
  const int32_t *srow = sub + (int64_t)hloc[P.hoff + i] * ld;
  const int jend = max(P.ldv, P.ldb);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < jend; j += gridDim.x * blockDim.x) {
    if (j < P.ldb)
      lbb[P.lbb_off + (int64_t)i * P.ldb + j] = j < P.b ? srow[hloc[P.hoff + j]] : kInf;
    if (j < P.ldv)
      lbv[P.lbv_off + (int64_t)i * P.ldv + j] =
          j < P.n ? srow[ed2st[colid[P.start + j]] - P.start] : kInf;
  }
}

}  // namespace fapk

namespace fapi {

namespace {

void free_bidx(fap_bidx *x) {
  if (!x) return;
  void *bufs[] = {x->parts,   x->in_off, x->in_arc, x->in1,      x->hloc,    x->inb_off, x->inb_pos_off,
                  x->inb_pos, x->outb_off, x->outb,  x->lbb,      x->lbv,     x->colid,
                  x->ctiles};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete x;
}

template <typename T>
int dev_alloc(int64_t count, T **d, size_t *bytes) {
  const size_t nb = (size_t)count * sizeof(T);
  hipError_t e = hipMalloc((void **)d, nb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *d = nullptr;
    return fail(FAP_E_NOMEM, "boundary index: device allocation of " + std::to_string(nb) +
                                 " bytes failed");
  }
  *bytes += nb;
  return FAP_OK;
}

// Copies h into a new device array (h must outlive the copy: the caller
// synchronizes s before h goes away).
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

constexpr int32_t kMaxIndexH = 1 << 18;  // boundary vertices (H ids): X rows of 1 MiB at most

}  // namespace

void bidx_release(fap_graph_dev *g) {
  free_bidx(g->bidx);
  g->bidx = nullptr;
  g->bidx_state = 0;
}

// Builds the index of g (partition attached).  Leaves g->bidx null (and
// returns FAP_OK) when the graph does not qualify: more than 2^18 boundary
// vertices, boundary vertices not first in their part's st2ed range, or
// duplicate arcs (the part matrix keeps the last duplicate, as the
// reference's does, which need not be the shortest).  Synchronous.
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
    P.ldb = (P.b + 3) & ~3;
    P.ldv = (P.n + 3) & ~3;
    P.lbb_off = lbb;
    lbb += (int64_t)P.b * P.ldb;
    P.lbv_off = lbv;
    lbv += (int64_t)P.b * P.ldv;
    pos += P.n;
  }
  if (hoff == 0 || hoff > kMaxIndexH) {
    delete x;
    return FAP_OK;
  }
  x->nbp = (int32_t)hoff;
  x->nb = 0;
  for (int32_t q = 1; q <= K; q++) x->nb += bcnt[q];
  {
    double wsum = 0;
    for (int64_t e = 0; e < g->E; e++) wsum += ed[e].y;
    const double mean_w = g->E ? wsum / (double)g->E : 1.0;
    x->delta = (int32_t)std::min<double>(std::max(1.0, mean_w * std::sqrt((double)V / K)),
                                         kInf / 4);
  }
  // H order: each part's boundary vertices grouped by the neighbouring part
  // of their first cut arc (either direction), st2ed order within a group
  std::vector<int32_t> firstnb(V, INT32_MAX);
  for (int32_t u = 0; u < V; u++)
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
      const int32_t v = ed[e].x;
      if (gid[v] == gid[u]) continue;
      firstnb[u] = std::min(firstnb[u], gid[v]);
      firstnb[v] = std::min(firstnb[v], gid[u]);
    }
  std::vector<int32_t> hid(V, -1), hloc((size_t)x->nbp, 0);
  x->boff.assign(K + 1, 0);
  x->h_hpos.assign(x->nb, 0);
  for (int32_t q = 0; q < K; q++) {
    const BidxPart &P = x->h_parts[q];
    std::vector<int32_t> ord(P.b);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
      return firstnb[st2ed[P.start + a]] < firstnb[st2ed[P.start + b]];
    });
    x->boff[q + 1] = x->boff[q] + P.b;
    for (int32_t h = 0; h < P.b; h++) {
      hid[st2ed[P.start + ord[h]]] = P.hoff + h;
      hloc[P.hoff + h] = ord[h];
      x->h_hpos[x->boff[q] + ord[h]] = h;
    }
  }
  // cut arcs into each H id; the borders: for each ordered pair of parts
  // (q, q2) with a cut arc q -> q2, the H positions of q2 it enters
  std::vector<int32_t> in_off(x->nbp + 1, 0);
  std::vector<int2> in_arc, in1;
  std::vector<int32_t> inb_off(K + 1, 0), inb_pos_off(1, 0), inb_pos, outb_off(K + 1, 0);
  std::vector<int2> outb;
  {
    std::vector<std::vector<int2>> lists(x->nbp);
    // per entered part q2: {source part q, H position of q2}
    std::vector<std::vector<std::pair<int32_t, int32_t>>> cuts(K);
    for (int32_t u = 0; u < V; u++)
      for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
        const int32_t v = ed[e].x;
        if (gid[v] == gid[u]) continue;
        lists[hid[v]].push_back(make_int2(hid[u], ed[e].y));
        cuts[gid[v] - 1].push_back({gid[u] - 1, hid[v] - x->h_parts[gid[v] - 1].hoff});
      }
    in1.assign(x->nbp, make_int2(-1, 0));
    for (int32_t h = 0; h < x->nbp; h++) {
      in_off[h + 1] = in_off[h] + (int32_t)lists[h].size();
      in_arc.insert(in_arc.end(), lists[h].begin(), lists[h].end());
      in1[h] = lists[h].size() == 1 ? lists[h][0] : make_int2(-1 - in_off[h], in_off[h + 1]);
    }
    std::vector<std::vector<int2>> outs(K);  // part q: {q2, slot of q in q2}
    for (int32_t q2 = 0; q2 < K; q2++) {
      std::vector<std::pair<int32_t, int32_t>> &c = cuts[q2];
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
      int32_t slot = 0;
      for (size_t i = 0; i < c.size(); slot++) {
        const int32_t q = c[i].first;
        while (i < c.size() && c[i].first == q) inb_pos.push_back(c[i++].second);
        inb_pos_off.push_back((int32_t)inb_pos.size());
        outs[q].push_back(make_int2(q2, slot));
      }
      inb_off[q2 + 1] = inb_off[q2] + slot;
      const int32_t kl = inb_pos_off.back() - inb_pos_off[inb_off[q2]];
      if (kl > kBidxKlMax) {  // a tile's border list would not fit in LDS
        delete x;
        return FAP_OK;
      }
    }
    for (int32_t q = 0; q < K; q++) {
      outb_off[q + 1] = outb_off[q] + (int32_t)outs[q].size();
      outb.insert(outb.end(), outs[q].begin(), outs[q].end());
    }
  }
  std::vector<int32_t> colid(st2ed);
  std::vector<int2> ctiles;
  for (int32_t q = 0; q < K; q++) {
    const BidxPart &P = x->h_parts[q];
    std::sort(colid.begin() + P.start, colid.begin() + P.start + P.n);
    for (int32_t j = 0; j < P.n; j += kExpandCols) ctiles.push_back(make_int2(q, j));
  }
  x->nctiles = (int32_t)ctiles.size();
  int rc = FAP_OK;
  if (rc == FAP_OK) rc = upload(x->h_parts, &x->parts, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(in_off, &x->in_off, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(in_arc, &x->in_arc, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(in1, &x->in1, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(hloc, &x->hloc, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(inb_off, &x->inb_off, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(inb_pos_off, &x->inb_pos_off, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(inb_pos, &x->inb_pos, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(outb_off, &x->outb_off, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(outb, &x->outb, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(colid, &x->colid, &x->bytes, s);
  if (rc == FAP_OK) rc = upload(ctiles, &x->ctiles, &x->bytes, s);
  // (every word of both is written by the extraction below)
  if (rc == FAP_OK) rc = dev_alloc(std::max<int64_t>(lbb, 1), &x->lbb, &x->bytes);
  if (rc == FAP_OK) rc = dev_alloc(std::max<int64_t>(lbv, 1), &x->lbv, &x->bytes);
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
      hipLaunchKernelGGL(fapk::bidx_extract_kernel,
                         dim3((std::max(P.ldv, P.ldb) + 255) / 256, P.b), dim3(256), 0, s,
                         subs[i], lds[i], P, x->hloc, x->colid, g->ed2st, x->lbb, x->lbv);
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

namespace {

// Device workspace of the block iterations (`bytes`) and its pinned host
// staging (`hbytes`, the leading part uploaded once), grown on demand.
int bbf_buffers(fap_ctx *ctx, size_t bytes, size_t hbytes, char **h, char **d) {
  if (ctx->host_bbf_bytes < hbytes) {
    if (ctx->host_bbf) (void)hipHostFree(ctx->host_bbf);
    ctx->host_bbf = nullptr;
    ctx->host_bbf_bytes = 0;
    FAP_HIP(hipHostMalloc(&ctx->host_bbf, hbytes, hipHostMallocDefault));
    ctx->host_bbf_bytes = hbytes;
  }
  FAP_TRY(ensure_ws(ctx, ctx->ws_bbf, bytes));
  *h = (char *)ctx->host_bbf;
  *d = (char *)ctx->ws_bbf.ptr;
  return FAP_OK;
}

}  // namespace

// Stage-1 rows for the boundary vertices of `count` parts (part indices
// qidx, X rows from rowoff[p], nsrc rows in all; d_rowptr: device row
// pointers of the output rows, V columns each, in the same order): the
// block iterations, then the expansion, all enqueued on s.  The host looks
// at the done word every kBatch iterations (one wait on s each time).
// Uses ctx->ws_bidx (nsrc x nbp int32) and ctx->ws_bbf / host_bbf.
int bidx_rows(fap_ctx *ctx, const fap_graph_dev *g, int32_t count, const int32_t *qidx,
              const int32_t *rowoff, int32_t nsrc, int32_t *const *d_rowptr, hipStream_t s) {
  const fap_bidx *x = g->bidx;
  if (!x) return fail(FAP_E_STATE, "no boundary index");
  if (nsrc <= 0) return FAP_OK;
  const int rows64 = (nsrc + 63) / 64;
  if (rows64 > 65535) return fail(FAP_E_ARG, "boundary rows: too many sources in one call");
  const int32_t K = x->K;
  const int64_t nblk = (int64_t)count * K;
  if (nblk > INT32_MAX / 2) return fail(FAP_E_ARG, "boundary rows: too many parts in one call");
  const int64_t ldx = x->nbp;  // a multiple of 4
  FAP_TRY(ensure_ws(ctx, ctx->ws_bidx, (size_t)nsrc * ldx * sizeof(int32_t)));
  int32_t *X = (int32_t *)ctx->ws_bidx.ptr;
  // workspace: control | block rows {first X row, rows} | source rows {part,
  // H position} | pending | marks | tiles
  // block-update tile: 64 rows x tcol columns (tcol = 64 or 128)
  const int tcol = ctx->opt_bidx_tile == 64 ? 64 : 128;
  int64_t rt_sum = 0, ct_sum = 0;
  for (int32_t p = 0; p < count; p++) rt_sum += (x->h_parts[qidx[p]].b + 63) / 64;
  for (int32_t q = 0; q < K; q++) ct_sum += (x->h_parts[q].b + tcol - 1) / tcol;
  const int64_t max_tiles = std::max<int64_t>(rt_sum * ct_sum, 1);
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t o_prow = up(sizeof(int32_t) * BC_WORDS);
  const size_t o_rowq = o_prow + up(sizeof(int2) * count);
  const size_t o_pend = o_rowq + up(sizeof(int2) * nsrc);
  const size_t o_mask = o_pend + up(sizeof(int32_t) * nblk);
  const size_t o_tile = o_mask + up(sizeof(uint32_t) * nblk);
  const size_t total = o_tile + sizeof(BidxTile) * (size_t)max_tiles;
  char *h = nullptr, *d = nullptr;
  FAP_TRY(bbf_buffers(ctx, total, o_tile, &h, &d));
  int32_t *hc = (int32_t *)h;
  std::fill(hc, hc + BC_WORDS, 0);
  // band: the graph's (mean arc weight x sqrt(V / K)) for 32 or more parts
  // a call, widened for fewer -- with little work per iteration, fewer
  // iterations pay more than the extra updates a wide band lets through
  // (grid, one part per call: 10.1 ms at 1x, 6.8 ms at 32x)
  const int64_t auto_delta = (int64_t)x->delta * std::max<int32_t>(1, 32 / std::max(count, 1));
  const int32_t delta = (int32_t)std::min<int64_t>(
      std::max<int64_t>(1, ctx->opt_bidx_delta > 0 ? ctx->opt_bidx_delta : auto_delta), kInf / 4);
  hc[BC_THR] = delta;
  int2 *hprow = (int2 *)(h + o_prow);
  int2 *hrowq = (int2 *)(h + o_rowq);
  int32_t *hpend = (int32_t *)(h + o_pend);
  std::fill(hpend, hpend + nblk, kInf);
  memset(h + o_mask, 0, sizeof(uint32_t) * nblk);
  for (int32_t p = 0; p < count; p++) {
    const int32_t q = qidx[p], b = x->h_parts[q].b;
    hprow[p] = make_int2(rowoff[p], b);
    for (int32_t i = 0; i < b; i++) hrowq[rowoff[p] + i] = make_int2(q, x->h_hpos[x->boff[q] + i]);
    if (b > 0) hpend[(size_t)p * K + q] = 0;  // each block starts as its own part's distances
  }
  FAP_HIP(hipMemcpyAsync(d, h, o_tile, hipMemcpyHostToDevice, s));
  int32_t *dctrl = (int32_t *)d;
  hipLaunchKernelGGL(fapk::bidx_block_init_kernel, dim3(nsrc), dim3(256), 0, s, X, ldx, x->nbp,
                     (const int2 *)(d + o_rowq), x->parts, x->lbb);
  FAP_CHECK_LAUNCH();
  // Bellman-Ford over the parts, block by block: a block (p, q2) is updated
  // after a block (p, q) with a cut arc q -> q2 changed, over the border
  // positions of q2 entered from such parts q.  Changed blocks wait in a
  // distance band order (delta-stepping over blocks): an iteration
  // propagates the pending blocks whose smallest changed value lies below
  // the current threshold, raised by delta when none does, so that most
  // blocks are propagated once their values are close to final.  Every
  // change is propagated eventually, so the result does not depend on the
  // band; values only decrease, so the iterations end.
  const int grid = (int)std::min<int64_t>(max_tiles, 8 * std::max(ctx->num_cus, 1));
  constexpr int kBatch = 16;
  for (int64_t it = 0;; it += kBatch) {
    if (it > ((int64_t)1 << 22)) return fail(FAP_E_STATE, "boundary iterations did not settle");
    for (int i = 0; i < kBatch; i++) {
      hipLaunchKernelGGL(fapk::bidx_schedule_kernel, dim3(1), dim3(1024), 0, s, dctrl,
                         (int32_t *)(d + o_pend), (uint32_t *)(d + o_mask), (BidxTile *)(d + o_tile), (const int2 *)(d + o_prow), count, K, delta, tcol,
                         x->parts, x->outb_off, x->outb);
      FAP_CHECK_LAUNCH();
      if (tcol == 64)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(fapk::bidx_block_kernel<4, 4>), dim3(grid), dim3(fapk::MP_THREADS), 0, s,
                           (const int32_t *)dctrl, (const BidxTile *)(d + o_tile), X, ldx,
                           x->parts, x->in1, x->in_arc, x->lbb, x->inb_off, x->inb_pos_off,
                           x->inb_pos, (int32_t *)(d + o_pend));
      else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(fapk::bidx_block_kernel<4, 8>), dim3(grid), dim3(fapk::MP_THREADS), 0, s,
                           (const int32_t *)dctrl, (const BidxTile *)(d + o_tile), X, ldx,
                           x->parts, x->in1, x->in_arc, x->lbb, x->inb_off, x->inb_pos_off,
                           x->inb_pos, (int32_t *)(d + o_pend));
      FAP_CHECK_LAUNCH();
    }
    FAP_HIP(hipMemcpyAsync(h, d, sizeof(int32_t) * BC_WORDS, hipMemcpyDeviceToHost, s));
    FAP_HIP(hipStreamSynchronize(s));
    if (hc[BC_DONE]) break;
  }
  hipLaunchKernelGGL(HIP_KERNEL_NAME(fapk::bidx_expand_kernel<kExpandTM, kExpandTN>),
                     dim3(x->nctiles, (nsrc + 16 * kExpandTM - 1) / (16 * kExpandTM)),
                     dim3(fapk::MP_THREADS), 0, s, X, ldx, nsrc, d_rowptr, x->ctiles, x->parts,
                     x->lbv, x->colid);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int32_t bidx_part(const fap_graph_dev *g, int32_t start, int32_t n, int32_t b) {
  const fap_bidx *x = g->bidx;
  if (!x || start < 0 || start >= g->V || n <= 0) return -1;
  const int32_t q = g->h_graph_id[g->h_st2ed[start]] - 1;
  if (q < 0 || q >= x->K) return -1;
  const BidxPart &P = x->h_parts[q];
  return (P.start == start && P.n == n && P.b == b) ? q : -1;
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
