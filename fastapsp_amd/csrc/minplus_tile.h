// minplus_tile.h -- register-blocked VALU min-plus tile for gfx950 (wave64).
//
// One workgroup of 256 threads (4 waves) computes a 128 x 128 (or 64 x 64) tile
//   C[i][j] = min(base, min_k A[i][k] + B[k][j]),  base = C[i][j] or INF.
// Min-plus is not a multiply-accumulate, so it runs on the VALU, not MFMA:
// every update is one v_add_u32, and two consecutive k are folded with one
// v_min3_i32, i.e. 1.5 VALU instructions per update, issued as interleaved
// (add, add, min3) triples so that adds of different waves can pair.
//
// Layout:
//   * waves form a 2 x 2 grid of sub-tiles; a lane owns a TM x TN
//     micro-tile (8 x 8 or 4 x 4): rows ly + 8r (r < TM, interleaved so that
//     the 8 distinct rows a wave reads from LDS fall into distinct banks) and
//     columns lx*4 + 32h + c (h < TN/4, c < 4: 16-byte groups, so each group
//     of 8 lanes reads 128 contiguous bytes of a B row).
//   * k is staged through LDS in chunks of 16: A as As[m][k] (row stride 20
//     words), B as Bs[k][n] (row stride BN + 4 words); global loads for
//     chunk c+1 are issued before chunk c is computed.
//   * out-of-range rows, columns and k read as INF, so ragged M, N, K need no
//     padding (INF + anything >= INF never wins a strict min).
#ifndef FAP_MINPLUS_TILE_H_
#define FAP_MINPLUS_TILE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace fapk {

constexpr int kInf = 0x3FFFFFFF;
constexpr int MP_BK = 16;
constexpr int MP_THREADS = 256;
constexpr int MP_AS_LD = MP_BK + 4;  // words per A row in LDS

// Flags describing what 16-byte vector accesses are legal for a launch.
enum : int {
  MP_VEC_A = 1,  // A base 16-byte aligned and lda % 4 == 0
  MP_VEC_B = 2,  // same for B
  MP_VEC_C = 4,  // same for C
};

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Tile geometry: each lane owns TM x TN outputs (TM, TN in {4, 8}); the 4
// waves form a 2 x 2 grid, so the workgroup tile is (16 TM) x (16 TN):
// 128 x 128 for 8 x 8, 64 x 64 for 4 x 4.
template <int TM, int TN>
struct MpTile {
  static constexpr int BM = 16 * TM;
  static constexpr int BN = 16 * TN;
  static constexpr int BS_LD = BN + 4;  // words per B row in LDS
  static constexpr int LDS_WORDS = BM * MP_AS_LD + MP_BK * BS_LD;
  static constexpr int A_TPR = MP_BK / TM;  // threads per A row when staging

  // Stage one 16-wide k chunk into registers: A row t / A_TPR, k
  // (t % A_TPR) * TM .. +TM; B k-row t / 16, columns (t % 16) * TN .. +TN.
  __device__ static __forceinline__ void load_chunk(const int32_t *A, int64_t lda,
                                                    const int32_t *B, int64_t ldb, int m_valid,
                                                    int n_valid, int K, int k0, int flags,
                                                    int ra[TM], int rb[TN]) {
    const int t = threadIdx.x;
    {
      const int row = t / A_TPR;
      const int kq = k0 + (t % A_TPR) * TM;
      const int32_t *p = A + (int64_t)row * lda + kq;
      if (row < m_valid && kq + TM <= K && (flags & MP_VEC_A)) {
#pragma unroll
        for (int q = 0; q < TM; q += 4) {
          const int4 v = *reinterpret_cast<const int4 *>(p + q);
          ra[q] = v.x; ra[q + 1] = v.y; ra[q + 2] = v.z; ra[q + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < TM; i++) ra[i] = (row < m_valid && kq + i < K) ? p[i] : kInf;
      }
    }
    {
      const int kr = k0 + (t >> 4);
      const int col = (t & 15) * TN;
      const int32_t *p = B + (int64_t)kr * ldb + col;
      if (kr < K && col + TN <= n_valid && (flags & MP_VEC_B)) {
#pragma unroll
        for (int q = 0; q < TN; q += 4) {
          const int4 v = *reinterpret_cast<const int4 *>(p + q);
          rb[q] = v.x; rb[q + 1] = v.y; rb[q + 2] = v.z; rb[q + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < TN; i++) rb[i] = (kr < K && col + i < n_valid) ? p[i] : kInf;
      }
    }
  }

  __device__ static __forceinline__ void store_chunk(int32_t *As, int32_t *Bs, const int ra[TM],
                                                     const int rb[TN]) {
    const int t = threadIdx.x;
    int32_t *pa = As + (t / A_TPR) * MP_AS_LD + (t % A_TPR) * TM;
#pragma unroll
    for (int q = 0; q < TM; q += 4)
      *reinterpret_cast<int4 *>(pa + q) = make_int4(ra[q], ra[q + 1], ra[q + 2], ra[q + 3]);
    int32_t *pb = Bs + (t >> 4) * BS_LD + (t & 15) * TN;
#pragma unroll
    for (int q = 0; q < TN; q += 4)
      *reinterpret_cast<int4 *>(pb + q) = make_int4(rb[q], rb[q + 1], rb[q + 2], rb[q + 3]);
  }

  // acc = min(acc, As (x) Bs) over the first kend (even) k of the staged
  // chunk, in this lane's TM x TN cells (rows row0 + 8r, columns
  // col0 + 32h + c).  One (add, add, min3) triple per two updates, in this
  // order: on gfx950 v_add_u32 issues from two waves of a SIMD in one
  // cycle; left to itself the compiler emits all adds of a k pair and then
  // all mins, so the waves sit in min-only stretches that cannot pair.
  // Interleaved, the 8192^3 product runs at 68 instead of 50 TOP/s
  // (measured ceiling of this mix: 76).  The asm is volatile only to keep
  // the triples in order; it touches registers alone.
  __device__ static __forceinline__ void compute(const int32_t *As, const int32_t *Bs, int kend,
                                                 int acc[TM][TN]) {
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int lx = lane & 7, ly = lane >> 3;
    const int row0 = (wave >> 1) * (8 * TM) + ly;
    const int col0 = (wave & 1) * (8 * TN) + lx * 4;
#pragma unroll 1
    for (int kk = 0; kk < kend; kk += 2) {
      int a0[TM], a1[TM], b0[TN], b1[TN];
#pragma unroll
      for (int r = 0; r < TM; r++) {
        const int2 v = *reinterpret_cast<const int2 *>(As + (row0 + 8 * r) * MP_AS_LD + kk);
        a0[r] = v.x;
        a1[r] = v.y;
      }
#pragma unroll
      for (int h = 0; h < TN / 4; h++) {
        const int4 u0 = *reinterpret_cast<const int4 *>(Bs + kk * BS_LD + col0 + 32 * h);
        const int4 u1 = *reinterpret_cast<const int4 *>(Bs + (kk + 1) * BS_LD + col0 + 32 * h);
        b0[4 * h + 0] = u0.x; b0[4 * h + 1] = u0.y; b0[4 * h + 2] = u0.z; b0[4 * h + 3] = u0.w;
        b1[4 * h + 0] = u1.x; b1[4 * h + 1] = u1.y; b1[4 * h + 2] = u1.z; b1[4 * h + 3] = u1.w;
      }
#pragma unroll
      for (int r = 0; r < TM; r++)
#pragma unroll
        for (int c = 0; c < TN; c++) {
          int t0, t1;
          asm volatile(
              "v_add_u32 %1, %3, %4\n\tv_add_u32 %2, %5, %6\n\tv_min3_i32 %0, %0, %1, %2"
              : "+v"(acc[r][c]), "=&v"(t0), "=&v"(t1)
              : "v"(a0[r]), "v"(b0[c]), "v"(a1[r]), "v"(b1[c]));
        }
    }
  }

  // Compute the tile.  A points at the tile's first row (k = 0), B at the
  // tile's first column (k = 0), C at the tile's corner; m_valid / n_valid
  // are the in-range rows / columns (<= BM / BN).  `lds` holds LDS_WORDS
  // ints.  When C aliases A or B (Floyd-Warshall panels), every read of the
  // aliased operand happens before the epilogue writes C, and the tile is
  // owned by this workgroup alone.
  __device__ static __forceinline__ void run(const int32_t *A, int64_t lda, const int32_t *B,
                                             int64_t ldb, int32_t *C, int64_t ldc, int m_valid,
                                             int n_valid, int K, bool accumulate, int flags,
                                             int32_t *lds) {
    int32_t *As = lds;
    int32_t *Bs = lds + BM * MP_AS_LD;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int lx = lane & 7, ly = lane >> 3;
    const int row0 = (wave >> 1) * (8 * TM) + ly;    // rows row0 + 8r
    const int col0 = (wave & 1) * (8 * TN) + lx * 4;  // cols col0 + 32h + c

    int acc[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; r++)
#pragma unroll
      for (int c = 0; c < TN; c++) acc[r][c] = kInf;

    int ra[TM], rb[TN];
    const int nchunks = (K + MP_BK - 1) / MP_BK;
    if (nchunks > 0) {
      load_chunk(A, lda, B, ldb, m_valid, n_valid, K, 0, flags, ra, rb);
      store_chunk(As, Bs, ra, rb);
    }
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
      const bool more = ch + 1 < nchunks;
      if (more) load_chunk(A, lda, B, ldb, m_valid, n_valid, K, (ch + 1) * MP_BK, flags, ra, rb);
      // the last chunk stops after its last valid k pair (K = 115 computes
      // 116 k, not 128); a k past K reads INF from LDS and changes nothing
      const int krem = K - ch * MP_BK;
      const int kend = krem >= MP_BK ? MP_BK : (krem + 1) & ~1;
      compute(As, Bs, kend, acc);
      __syncthreads();
      if (more) {
        store_chunk(As, Bs, ra, rb);
        __syncthreads();
      }
    }

    // epilogue.  Accumulating into a full tile, the C pieces of RG rows are
    // all loaded before any of them is stored: stores to C may alias later
    // loads as far as the compiler knows (the row stride is a run-time
    // value), so loads interleaved with stores would each wait for the
    // previous piece's round trip to memory.
    if (accumulate && m_valid == BM && n_valid == BN && (flags & MP_VEC_C)) {
      constexpr int HN = TN / 4;            // 16-byte pieces per row and lane
      constexpr int RG = TM >= 8 ? 2 : 4;   // rows per round trip
#pragma unroll
      for (int r0 = 0; r0 < TM; r0 += RG) {
        int4 o[RG][HN];
#pragma unroll
        for (int rr = 0; rr < RG; rr++)
#pragma unroll
          for (int h = 0; h < HN; h++)
            o[rr][h] = *reinterpret_cast<const int4 *>(C + (int64_t)(row0 + 8 * (r0 + rr)) * ldc +
                                                       col0 + 32 * h);
#pragma unroll
        for (int rr = 0; rr < RG; rr++)
#pragma unroll
          for (int h = 0; h < HN; h++) {
            const int r = r0 + rr;
            *reinterpret_cast<int4 *>(C + (int64_t)(row0 + 8 * r) * ldc + col0 + 32 * h) =
                make_int4(imin(acc[r][4 * h + 0], o[rr][h].x), imin(acc[r][4 * h + 1], o[rr][h].y),
                          imin(acc[r][4 * h + 2], o[rr][h].z), imin(acc[r][4 * h + 3], o[rr][h].w));
          }
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < TM; r++) {
      const int row = row0 + 8 * r;
      if (row >= m_valid) continue;
      int32_t *crow = C + (int64_t)row * ldc;
#pragma unroll
      for (int h = 0; h < TN / 4; h++) {
        const int col = col0 + 32 * h;
        if ((flags & MP_VEC_C) && col + 4 <= n_valid) {
          int4 *cp = reinterpret_cast<int4 *>(crow + col);
          int4 v = make_int4(acc[r][4 * h + 0], acc[r][4 * h + 1], acc[r][4 * h + 2],
                             acc[r][4 * h + 3]);
          if (accumulate) {
            const int4 o = *cp;
            v.x = imin(v.x, o.x);
            v.y = imin(v.y, o.y);
            v.z = imin(v.z, o.z);
            v.w = imin(v.w, o.w);
          }
          *cp = v;
        } else {
#pragma unroll
          for (int c = 0; c < 4; c++) {
            if (col + c < n_valid) {
              int v = acc[r][4 * h + c];
              if (accumulate) v = imin(v, crow[col + c]);
              crow[col + c] = v;
            }
          }
        }
      }
    }
  }
};

using MpBig = MpTile<8, 8>;    // 128 x 128
using MpSmall = MpTile<4, 4>;  // 64 x 64

// Floyd-Warshall closure of one diagonal block of up to 64 x 64 (kb valid
// rows and columns; the rest read as INF) with 256 threads, in LDS (`D`
// holds BLK_LDS_WORDS ints, row stride BLK_LD): thread t owns the 1 x 4
// strips at rows (t >> 4) + 16q, q < 4, columns (t & 15) * 4.  Row k and
// column k do not change during pivot k (the diagonal is >= 0), so every
// thread updates in place; one barrier per pivot.
// Closure of a 64 x 64 sub-block with 1024 threads: a thread owns the 1 x 4
// strip at row t >> 4, columns (t & 15) * 4 .. +3 of the sub-block.
// The sub-block is at (o, o) of D (row stride LD), pivots o .. o + nv - 1.  Row k
// and column k do not change during step k (the diagonal is >= 0), so every
// thread can update its strip in place; one barrier per pivot.
template <int LD>
__device__ __forceinline__ void lds_fw64(int32_t *D, int o, int nv) {
  const int t = threadIdx.x;
  const int i = o + (t >> 4), j = o + ((t & 15) << 2);
  int4 *cp = reinterpret_cast<int4 *>(D + i * LD + j);
  for (int k = o; k < o + nv; k++) {
    const int a = D[i * LD + k];
    const int4 b = *reinterpret_cast<const int4 *>(D + k * LD + j);
    int4 c = *cp;
    c.x = imin(c.x, a + b.x);
    c.y = imin(c.y, a + b.y);
    c.z = imin(c.z, a + b.z);
    c.w = imin(c.w, a + b.w);
    *cp = c;
    __syncthreads();
  }
}


constexpr int BLK_LD = 64 + 4;
constexpr int BLK_LDS_WORDS = 64 * BLK_LD;

__device__ __forceinline__ void close_block64(int32_t *blk, int64_t ld, int kb, int32_t *D) {
  const int t = threadIdx.x;
  for (int idx = t; idx < 64 * 64; idx += MP_THREADS) {
    const int i = idx >> 6, j = idx & 63;
    D[i * BLK_LD + j] = (i < kb && j < kb) ? blk[(int64_t)i * ld + j] : kInf;
  }
  __syncthreads();
  const int r0 = t >> 4, j = (t & 15) << 2;
  for (int k = 0; k < kb; k++) {
    const int4 b = *reinterpret_cast<const int4 *>(D + k * BLK_LD + j);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = r0 + 16 * q;
      const int a = D[i * BLK_LD + k];
      int4 *cp = reinterpret_cast<int4 *>(D + i * BLK_LD + j);
      int4 c = *cp;
      c.x = imin(c.x, a + b.x);
      c.y = imin(c.y, a + b.y);
      c.z = imin(c.z, a + b.z);
      c.w = imin(c.w, a + b.w);
      *cp = c;
    }
    __syncthreads();
  }
  for (int idx = t; idx < 64 * 64; idx += MP_THREADS) {
    const int i = idx >> 6, j2 = idx & 63;
    if (i < kb && j2 < kb) blk[(int64_t)i * ld + j2] = D[i * BLK_LD + j2];
  }
}
constexpr int MP_BM = MpBig::BM;
constexpr int MP_BN = MpBig::BN;

}  // namespace fapk

#endif  // FAP_MINPLUS_TILE_H_
