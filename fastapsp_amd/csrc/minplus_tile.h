// minplus_tile.h -- register-blocked VALU min-plus tile for gfx950 (wave64).
//
// One workgroup of 256 threads (4 waves) computes a 128 x 128 tile
//   C[i][j] = min(base, min_k A[i][k] + B[k][j]),  base = C[i][j] or INF.
// Min-plus is not a multiply-accumulate, so it runs on the VALU, not MFMA:
// every update is one v_add_u32, and two consecutive k are folded with one
// v_min3_i32, i.e. 1.5 VALU instructions per update.
//
// Layout:
//   * waves form a 2 x 2 grid of 64 x 64 sub-tiles; a lane owns an 8 x 8
//     micro-tile: rows ly + 8r (r < 8, interleaved so that the 8 distinct rows
//     a wave reads from LDS fall into distinct banks) and columns
//     lx*4 + 32h + c (h < 2, c < 4: two 16-byte groups, so each group of 8
//     lanes reads 128 contiguous bytes of a B row).
//   * k is staged through LDS in chunks of 16: A as As[m][k] (row stride 20
//     words), B as Bs[k][n] (row stride 132 words); global loads for chunk
//     c+1 are issued before chunk c is computed.
//   * out-of-range rows, columns and k read as INF, so ragged M, N, K need no
//     padding (INF + anything >= INF never wins a strict min).
#ifndef FAP_MINPLUS_TILE_H_
#define FAP_MINPLUS_TILE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fapk {

constexpr int kInf = 0x3FFFFFFF;
constexpr int MP_BM = 128;
constexpr int MP_BN = 128;
constexpr int MP_BK = 16;
constexpr int MP_THREADS = 256;
constexpr int MP_AS_LD = MP_BK + 4;   // words per A row in LDS
constexpr int MP_BS_LD = MP_BN + 4;   // words per B row in LDS
constexpr int MP_LDS_WORDS = MP_BM * MP_AS_LD + MP_BK * MP_BS_LD;

// Flags describing what 16-byte vector accesses are legal for a launch.
enum : int {
  MP_VEC_A = 1,  // A base 16-byte aligned and lda % 4 == 0
  MP_VEC_B = 2,  // same for B
  MP_VEC_C = 4,  // same for C
};

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Stage one 16-wide k chunk of A (tile rows) and B (tile columns) into
// registers.  Thread t: A row t>>1, k (t&1)*8..+8; B k-row t>>4, cols (t&15)*8..+8.
__device__ __forceinline__ void mp_load_chunk(const int32_t *__restrict__ A, int64_t lda,
                                              const int32_t *__restrict__ B, int64_t ldb,
                                              int m_valid, int n_valid, int K, int k0,
                                              int flags, int ra[8], int rb[8]) {
  const int t = threadIdx.x;
  {
    const int row = t >> 1;
    const int kq = k0 + ((t & 1) << 3);
    const int32_t *p = A + (int64_t)row * lda + kq;
    if (row < m_valid && kq + 8 <= K && (flags & MP_VEC_A)) {
      int4 v0 = *reinterpret_cast<const int4 *>(p);
      int4 v1 = *reinterpret_cast<const int4 *>(p + 4);
      ra[0] = v0.x; ra[1] = v0.y; ra[2] = v0.z; ra[3] = v0.w;
      ra[4] = v1.x; ra[5] = v1.y; ra[6] = v1.z; ra[7] = v1.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) ra[i] = (row < m_valid && kq + i < K) ? p[i] : kInf;
    }
  }
  {
    const int kr = k0 + (t >> 4);
    const int col = (t & 15) << 3;
    const int32_t *p = B + (int64_t)kr * ldb + col;
    if (kr < K && col + 8 <= n_valid && (flags & MP_VEC_B)) {
      int4 v0 = *reinterpret_cast<const int4 *>(p);
      int4 v1 = *reinterpret_cast<const int4 *>(p + 4);
      rb[0] = v0.x; rb[1] = v0.y; rb[2] = v0.z; rb[3] = v0.w;
      rb[4] = v1.x; rb[5] = v1.y; rb[6] = v1.z; rb[7] = v1.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) rb[i] = (kr < K && col + i < n_valid) ? p[i] : kInf;
    }
  }
}

__device__ __forceinline__ void mp_store_chunk(int32_t *As, int32_t *Bs, const int ra[8],
                                               const int rb[8]) {
  const int t = threadIdx.x;
  int32_t *pa = As + (t >> 1) * MP_AS_LD + ((t & 1) << 3);
  *reinterpret_cast<int4 *>(pa) = make_int4(ra[0], ra[1], ra[2], ra[3]);
  *reinterpret_cast<int4 *>(pa + 4) = make_int4(ra[4], ra[5], ra[6], ra[7]);
  int32_t *pb = Bs + (t >> 4) * MP_BS_LD + ((t & 15) << 3);
  *reinterpret_cast<int4 *>(pb) = make_int4(rb[0], rb[1], rb[2], rb[3]);
  *reinterpret_cast<int4 *>(pb + 4) = make_int4(rb[4], rb[5], rb[6], rb[7]);
}

// Compute the tile.  A points at the tile's first row (k = 0), B at the
// tile's first column (k = 0), C at the tile's corner.  m_valid / n_valid are
// the in-range rows / columns of this tile (<= 128).  `lds` holds
// MP_LDS_WORDS ints.  When C aliases A or B (Floyd-Warshall panels), every
// read of the aliased operand happens before the epilogue writes C, and the
// tile is owned by this workgroup alone.
__device__ __forceinline__ void minplus_tile(const int32_t *A, int64_t lda, const int32_t *B,
                                             int64_t ldb, int32_t *C, int64_t ldc, int m_valid,
                                             int n_valid, int K, bool accumulate, int flags,
                                             int32_t *lds) {
  int32_t *As = lds;
  int32_t *Bs = lds + MP_BM * MP_AS_LD;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int lx = lane & 7, ly = lane >> 3;
  const int row0 = (wave >> 1) * 64 + ly;        // rows row0 + 8r
  const int col0 = (wave & 1) * 64 + lx * 4;     // cols col0 + 32h + c

  int acc[8][8];
#pragma unroll
  for (int r = 0; r < 8; r++)
#pragma unroll
    for (int c = 0; c < 8; c++) acc[r][c] = kInf;

  int ra[8], rb[8];
  const int nchunks = (K + MP_BK - 1) / MP_BK;
  if (nchunks > 0) {
    mp_load_chunk(A, lda, B, ldb, m_valid, n_valid, K, 0, flags, ra, rb);
    mp_store_chunk(As, Bs, ra, rb);
  }
  __syncthreads();
  for (int ch = 0; ch < nchunks; ch++) {
    const bool more = ch + 1 < nchunks;
    if (more) mp_load_chunk(A, lda, B, ldb, m_valid, n_valid, K, (ch + 1) * MP_BK, flags, ra, rb);
#pragma unroll 1
    for (int kk = 0; kk < MP_BK; kk += 2) {
      int a0[8], a1[8], b0[8], b1[8];
#pragma unroll
      for (int r = 0; r < 8; r++) {
        int2 v = *reinterpret_cast<const int2 *>(As + (row0 + 8 * r) * MP_AS_LD + kk);
        a0[r] = v.x;
        a1[r] = v.y;
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        int4 u0 = *reinterpret_cast<const int4 *>(Bs + kk * MP_BS_LD + col0 + 32 * h);
        int4 u1 = *reinterpret_cast<const int4 *>(Bs + (kk + 1) * MP_BS_LD + col0 + 32 * h);
        b0[4 * h + 0] = u0.x; b0[4 * h + 1] = u0.y; b0[4 * h + 2] = u0.z; b0[4 * h + 3] = u0.w;
        b1[4 * h + 0] = u1.x; b1[4 * h + 1] = u1.y; b1[4 * h + 2] = u1.z; b1[4 * h + 3] = u1.w;
      }
#pragma unroll
      for (int r = 0; r < 8; r++)
#pragma unroll
        for (int c = 0; c < 8; c++)
          acc[r][c] = imin(acc[r][c], imin(a0[r] + b0[c], a1[r] + b1[c]));
    }
    __syncthreads();
    if (more) {
      mp_store_chunk(As, Bs, ra, rb);
      __syncthreads();
    }
  }

  // epilogue
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int row = row0 + 8 * r;
    if (row >= m_valid) continue;
    int32_t *crow = C + (int64_t)row * ldc;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int col = col0 + 32 * h;
      if ((flags & MP_VEC_C) && col + 4 <= n_valid) {
        int4 *cp = reinterpret_cast<int4 *>(crow + col);
        int4 v = make_int4(acc[r][4 * h + 0], acc[r][4 * h + 1], acc[r][4 * h + 2],
                           acc[r][4 * h + 3]);
        if (accumulate) {
          int4 o = *cp;
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

// Map a linear workgroup id to a tile index so that each XCD (workgroups are
// dealt round-robin over 8 XCDs) works on a contiguous range of tiles, which
// then share operand panels in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
  const int per = total >> 3;
  if (bid >= (per << 3)) return bid;
  return (bid & 7) * per + (bid >> 3);
}

}  // namespace fapk

#endif  // FAP_MINPLUS_TILE_H_
