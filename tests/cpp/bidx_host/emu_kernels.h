// CPU emulations of bindex.hip's kernels for the host check: each follows
// its kernel's indexing (tile maps, operand offsets, bounds), so that
// AddressSanitizer sees every access the device code would make, and runs
// the tiles one after another.  Results are min-plus exact like the
// kernels'; the GPU tests check the kernels themselves.
#pragma once
#include <algorithm>
#include <cassert>
namespace fapk {
constexpr int MP_THREADS = 256;
constexpr int MP_BK = 16;
using fapi::kInf;
static long g_tiles = 0, g_launches = 0, g_max_tiles = 0;

inline void bidx_block_init_kernel(dim3 g, dim3 b, int32_t *X, int64_t ldx, int32_t nbp,
                                   const int2 *rowq, const BidxPart *parts, const int32_t *lbb) {
  for (unsigned r = 0; r < g.x; r++) {
    const int2 qi = rowq[r];
    const BidxPart P = parts[qi.x];
    int32_t *xr = X + (int64_t)r * ldx;
    for (int c = 0; c < nbp; c++) {
      const int j = c - P.hoff;
      xr[c] = (j >= 0 && j < P.b) ? lbb[P.lbb_off + (int64_t)qi.y * P.ldb + j] : kInf;
    }
  }
}

static long g_kwork = 0;
inline void bidx_schedule_kernel(dim3 g, dim3 b, int32_t *ctrl, int32_t *pend, uint32_t *mask,
                                 BidxTile *tiles, const int2 *prow,
                                 int32_t count, int32_t K, int32_t delta, int32_t tcol, const BidxPart *parts,
                                 const int32_t *outb_off, const int2 *outb) {
  assert(g.x == 1 && b.x == 1024);
  const int nblk = count * K;
  int m = kInf;
  for (int i = 0; i < nblk; i++) m = std::min(m, pend[i]);
  int thr = ctrl[BC_THR];
  if (m < kInf && m >= thr) thr = (int)(((int64_t)m / delta + 1) * delta);
  ctrl[BC_THR] = thr;
  ctrl[BC_DONE] = m < kInf ? 0 : 1;
  if (m >= kInf) { ctrl[BC_NTILES] = 0; return; }
  for (int i = 0; i < nblk; i++) {
    if (pend[i] >= thr) continue;
    pend[i] = kInf;
    const int p = i / K, q = i - p * K;
    for (int e = outb_off[q]; e < outb_off[q + 1]; e++) {
      const int2 o = outb[e];
      assert(o.x >= 0 && o.x < K);
      if (parts[o.x].b > 0) mask[p * K + o.x] |= 1u << std::min(o.y, 31);
    }
  }
  int ntl = 0;
  for (int i = 0; i < nblk; i++) {
    const uint32_t mm = mask[i];
    if (!mm) continue;
    mask[i] = 0;
    const int p = i / K, q2 = i - p * K;
    const int2 pr = prow[p];
    const int rt = (pr.y + 63) >> 6, ct = (parts[q2].b + tcol - 1) / tcol;
    const int base = ntl;
    ntl += rt * ct;
    for (int tr = 0; tr < rt; tr++)
      for (int tc = 0; tc < ct; tc++) tiles[base + tr * ct + tc] = BidxTile{pr.x, pr.y, q2, tr, tc, i, mm, 0};
  }
  ctrl[BC_NTILES] = ntl;
}

template <int TM, int TN>
inline void bidx_block_kernel(dim3 g, dim3 b, const int32_t *ctrl, const BidxTile *tiles, int32_t *X,
                              int64_t ldx, const BidxPart *parts, const int2 *in1,
                              const int2 *in_arc, const int32_t *lbb, const int32_t *inb_off,
                              const int32_t *inb_pos_off, const int32_t *inb_pos, int32_t *pend) {
  assert(b.x == 256 && g.y == 1);
  const int ntiles = ctrl[BC_NTILES];
  if (ntiles == 0) return;
  g_launches++;
  g_tiles += ntiles;
  g_max_tiles = std::max<long>(g_max_tiles, ntiles);
  constexpr int TR = 16 * TM, TC = 16 * TN, ATPR = 16 / TM;
  static int A[TR][16], Bm[16][TC];
  static int klist[kBidxKlMax];
  for (int ti = 0; ti < ntiles; ti++) {
    const BidxTile T = tiles[ti];
    const BidxPart Q = parts[T.q];
    int K = 0;
    const int s0 = inb_off[T.q], s1 = inb_off[T.q + 1];
    for (int sl = s0; sl < s1; sl++) {
      const int si = sl - s0;
      if (!((T.slots >> std::min(si, 31)) & 1u)) continue;
      const int p0 = inb_pos_off[sl], len = inb_pos_off[sl + 1] - p0;
      assert(K + len <= kBidxKlMax);
      for (int i = 0; i < len; i++) klist[K + i] = inb_pos[p0 + i];
      K += len;
    }
    assert(K > 0);
    g_kwork += K;
    const int r0 = T.tr * TR, c0 = T.tc * TC;
    const int m_valid = std::min(TR, T.rows - r0), n_valid = std::min(TC, Q.b - c0);
    assert(m_valid > 0 && n_valid > 0);
    int32_t *Xt = X + (int64_t)(T.row0 + r0) * ldx;
    const int32_t *Bp = lbb + Q.lbb_off + c0;
    static int acc[TR][TC];
    for (int i = 0; i < TR; i++) for (int j = 0; j < TC; j++) acc[i][j] = kInf;
    const int nch = (K + MP_BK - 1) / MP_BK;
    for (int ch = 0; ch < nch; ch++) {
      const int k0 = ch * MP_BK;
      for (int t = 0; t < 256; t++) {
        const int arow = t / ATPR, akq = (t % ATPR) * TM;
        const int32_t *xa = Xt + (int64_t)arow * ldx;
        for (int j = 0; j < TM; j++) {
          const int k = k0 + akq + j;
          int v = kInf;
          if (arow < m_valid && k < K) {
            assert(klist[k] >= 0 && klist[k] < Q.b);
            const int2 f = in1[Q.hoff + klist[k]];
            if (f.x >= 0) v = std::min(kInf, xa[f.x] + f.y);
            else
              for (int e = -1 - f.x; e < f.y; e++) {
                const int2 a = in_arc[e];
                v = std::min(v, xa[a.x] + a.y);
              }
          }
          A[arow][akq + j] = v;
        }
        const int kr = k0 + (t >> 4), col = (t & 15) * TN;
        const int32_t *pb = Bp + (int64_t)(kr < K ? klist[kr] : 0) * Q.ldb + col;
        for (int j = 0; j < TN; j++) {
          int v;
          if (kr < K && col + TN <= n_valid) { assert(((uintptr_t)pb & 15) == 0); v = pb[j]; }
          else v = (kr < K && col + j < n_valid) ? pb[j] : kInf;
          Bm[t >> 4][col + j] = v;
        }
      }
      const int krem = K - k0;
      const int kend = krem >= MP_BK ? MP_BK : (krem + 1) & ~1;
      for (int i = 0; i < TR; i++)
        for (int k = 0; k < kend; k++) {
          const int a = A[i][k];
          for (int j = 0; j < TC; j++) acc[i][j] = std::min(acc[i][j], a + Bm[k][j]);
        }
    }
    int lowest = kInf;
    for (int row = 0; row < m_valid; row++) {
      int32_t *cr = Xt + (int64_t)row * ldx + Q.hoff + c0;
      for (int c = 0; c < n_valid; c++)
        if (acc[row][c] < cr[c]) { cr[c] = acc[row][c]; lowest = std::min(lowest, acc[row][c]); }
    }
    if (lowest < kInf) pend[T.blk] = std::min(pend[T.blk], lowest);
  }
}

template <int TM, int TN>
inline void bidx_expand_kernel(dim3 g, dim3 b, const int32_t *X, int64_t ldx, int32_t nsrc,
                               int32_t *const *rowptr, const int2 *ctiles, const BidxPart *parts,
                               const int32_t *lbv, const int32_t *colid) {
  constexpr int TR = 16 * TM, TC = 16 * TN;
  for (unsigned bx = 0; bx < g.x; bx++)
    for (unsigned by = 0; by < g.y; by++) {
      const int2 ct = ctiles[bx];
      const BidxPart P = parts[ct.x];
      const int j0 = ct.y, r0 = by * TR;
      const int m_valid = std::min(TR, nsrc - r0), n_valid = std::min(TC, P.n - j0), K = P.b;
      assert(m_valid > 0 && n_valid > 0);
      const int32_t *A = X + (int64_t)r0 * ldx + P.hoff;
      const int32_t *B = lbv + P.lbv_off + j0;
      assert(((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0);
      for (int row = 0; row < m_valid; row++) {
        int32_t *out = rowptr[r0 + row];
        for (int c = 0; c < n_valid; c++) {
          int v = kInf;
          for (int k = 0; k < K; k++) v = std::min(v, A[(int64_t)row * ldx + k] + B[(int64_t)k * P.ldv + c]);
          out[colid[P.start + j0 + c]] = v;
        }
      }
    }
}

inline void bidx_extract_kernel(dim3 g, dim3 b, const int32_t *sub, int64_t ld, BidxPart P,
                                const int32_t *hloc, const int32_t *colid, const int32_t *ed2st, int32_t *lbb,
                                int32_t *lbv) {
  for (unsigned i = 0; i < g.y; i++) {
    const int32_t *srow = sub + (int64_t)hloc[P.hoff + i] * ld;
    for (int j = 0; j < (int)(g.x * b.x) && j < std::max(P.ldv, P.ldb); j++) {
      if (j < P.ldb) lbb[P.lbb_off + (int64_t)i * P.ldb + j] = j < P.b ? srow[hloc[P.hoff + j]] : kInf;
      if (j < P.ldv)
        lbv[P.lbv_off + (int64_t)i * P.ldv + j] =
            j < P.n ? srow[ed2st[colid[P.start + j]] - P.start] : kInf;
    }
  }
}
}  // namespace fapk
