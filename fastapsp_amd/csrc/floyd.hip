// floyd.hip -- blocked Floyd-Warshall closure on the VALU, in place.
//
// Replaces the reference's stage-2 operator floyd_path (fap/kernel/floyd.h:
// 150-158; CPU semantics floyd_CPU_path :38-49: k-i-j order, strict relax).
// The closure of a non-negative matrix is unique, so the blocked order gives
// the same distances as the k-i-j loop.  With block size T (64 or 128) and
// nb = ceil(n / T) blocks, round r (pivot block r) runs:
//   1. diagonal tile (r, r): closed in one workgroup, in LDS;
//   2. row panel (r, J) = min(C, D_rr (x) C) and column panel
//      (I, r) = min(C, C (x) D_rr) for all I, J != r: one launch;
//   3. every other tile (I, J) = min(C, D_Ir (x) D_rJ): one launch.
// Phases 2 and 3 reuse the register-blocked min-plus tile (minplus_tile.h)
// with T x T tiles.  T = 128 halves the number of rounds and, with 8 x 8
// cells per lane, the LDS bytes per update (phase 3 is bound by LDS reads
// at 4 x 4 cells); T = 64 gives 4x more tiles per launch and a cheap
// diagonal phase.  128 is the default for large n, 64 below (launch_fw).
#include "common.h"
#include "minplus_tile.h"

namespace fapk {

constexpr int FW_DIAG_THREADS = 1024;

// In-LDS helpers for the diagonal tile (1024 threads).  The tile lives in
// D[T][T + 4] (rows 16-byte aligned); lds_fw64 (minplus_tile.h) closes a
// 64 x 64 sub-block of it.

// C = min(C, A (x) B) for 64 x 64 sub-blocks given by their corners
// (row, col) in D, K = 64.  All reads happen before the barrier that
// precedes the writes, so C may alias A or B.
template <int LD>
__device__ __forceinline__ void lds_mp64(int32_t *D, int ci, int cj, int ai, int ak, int bk,
                                         int bj) {
  const int t = threadIdx.x;
  const int r = t >> 4, c0 = (t & 15) << 2;
  int4 acc = *reinterpret_cast<const int4 *>(D + (ci + r) * LD + cj + c0);
  const int32_t *arow = D + (ai + r) * LD + ak;
#pragma unroll 4
  for (int k = 0; k < 64; k += 2) {
    const int2 a = *reinterpret_cast<const int2 *>(arow + k);
    const int4 b0 = *reinterpret_cast<const int4 *>(D + (bk + k) * LD + bj + c0);
    const int4 b1 = *reinterpret_cast<const int4 *>(D + (bk + k + 1) * LD + bj + c0);
    acc.x = imin(acc.x, imin(a.x + b0.x, a.y + b1.x));
    acc.y = imin(acc.y, imin(a.x + b0.y, a.y + b1.y));
    acc.z = imin(acc.z, imin(a.x + b0.z, a.y + b1.z));
    acc.w = imin(acc.w, imin(a.x + b0.w, a.y + b1.w));
  }
  __syncthreads();
  *reinterpret_cast<int4 *>(D + (ci + r) * LD + cj + c0) = acc;
  __syncthreads();
}

// Phase 1 for T = 64 (direct) or T = 128 (2 x 2 recursion on 64-blocks:
// close D00; panels D01, D10; D11 += D10 D01; close D11; panels D10, D01;
// D00 += D01 D10 -- the blocked algorithm itself, one level down, so most
// of the work runs as min-plus products with two k per v_min3).
template <int T>
__global__ __launch_bounds__(FW_DIAG_THREADS) void fw_diag_kernel(int32_t *d, int n,
                                                                  int64_t ld, int r) {
  constexpr int LD = T + 4;
  __shared__ __attribute__((aligned(16))) int32_t D[T * LD];
  const int base = r * T;
  const int kb = imin(T, n - base);
  int32_t *blk = d + (int64_t)base * ld + base;
  for (int idx = threadIdx.x; idx < T * T; idx += FW_DIAG_THREADS) {
    const int i = idx / T, j = idx - (idx / T) * T;
    D[i * LD + j] = (i < kb && j < kb) ? blk[(int64_t)i * ld + j] : kInf;
  }
  __syncthreads();
  const int n0 = imin(64, kb);
  lds_fw64<LD>(D, 0, n0);
  if (T == 128 && kb > 64) {
    const int n1 = kb - 64;
    lds_mp64<LD>(D, 0, 64, 0, 0, 0, 64);    // D01 = min(D01, D00 D01)
    lds_mp64<LD>(D, 64, 0, 64, 0, 0, 0);    // D10 = min(D10, D10 D00)
    lds_mp64<LD>(D, 64, 64, 64, 0, 0, 64);  // D11 = min(D11, D10 D01)
    lds_fw64<LD>(D, 64, n1);
    lds_mp64<LD>(D, 64, 0, 64, 64, 64, 0);  // D10 = min(D10, D11 D10)
    lds_mp64<LD>(D, 0, 64, 0, 64, 64, 64);  // D01 = min(D01, D01 D11)
    lds_mp64<LD>(D, 0, 0, 0, 64, 64, 0);    // D00 = min(D00, D01 D10)
  }
  for (int idx = threadIdx.x; idx < T * T; idx += FW_DIAG_THREADS) {
    const int i = idx / T, j = idx - (idx / T) * T;
    if (i < kb && j < kb) blk[(int64_t)i * ld + j] = D[i * LD + j];
  }
}

// Phase 2: workgroups [0, nb-1) take row-panel tiles (r, J), the next nb-1
// take column-panel tiles (I, r).  Each tile is owned by one workgroup and
// aliases one operand; the tile reads all operands before writing C.
template <int TM>
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void fw_panel_kernel(
    int32_t *d, int n, int64_t ld, int r, int nb, int flags) {
  using Tile = MpTile<TM, TM>;
  constexpr int T = Tile::BM;
  __shared__ __attribute__((aligned(16))) int32_t lds[Tile::LDS_WORDS];
  const int base = r * T;
  const int kb = imin(T, n - base);
  const int32_t *diag = d + (int64_t)base * ld + base;
  int bid = blockIdx.x;
  if (bid < nb - 1) {
    const int J = bid < r ? bid : bid + 1;
    int32_t *c = d + (int64_t)base * ld + (int64_t)J * T;
    Tile::run(diag, ld, c, ld, c, ld, kb, imin(T, n - J * T), kb, true, flags, lds);
  } else {
    bid -= nb - 1;
    const int I = bid < r ? bid : bid + 1;
    int32_t *c = d + (int64_t)I * T * ld + base;
    Tile::run(c, ld, diag, ld, c, ld, imin(T, n - I * T), kb, kb, true, flags, lds);
  }
}

// Closure of a 128 x 128 diagonal block (kb <= 128 valid rows/columns) by
// one 256-thread workgroup, in global memory: the 2 x 2 recursion of
// fw_diag_kernel<128> with each 64-block closed in LDS (close_block64) and
// each 64 x 64 product done by the product tile (C may alias A or B: the
// tile reads its operands before it writes).  `lds` holds BLK_LDS_WORDS.
__device__ __forceinline__ void close_block128(int32_t *blk, int64_t ld, int kb, int flags,
                                               int32_t *lds) {
  const int n0 = imin(64, kb), n1 = kb - 64;
  close_block64(blk, ld, n0, lds);
  __syncthreads();
  if (n1 <= 0) return;
  int32_t *d00 = blk, *d01 = blk + 64, *d10 = blk + 64 * ld, *d11 = blk + 64 * ld + 64;
  MpSmall::run(d00, ld, d01, ld, d01, ld, n0, n1, n0, true, flags, lds);  // D01 (+)= D00 D01
  __syncthreads();
  MpSmall::run(d10, ld, d00, ld, d10, ld, n1, n0, n0, true, flags, lds);  // D10 (+)= D10 D00
  __syncthreads();
  MpSmall::run(d10, ld, d01, ld, d11, ld, n1, n1, n0, true, flags, lds);  // D11 (+)= D10 D01
  __syncthreads();
  close_block64(d11, ld, n1, lds);
  __syncthreads();
  MpSmall::run(d11, ld, d10, ld, d10, ld, n1, n0, n1, true, flags, lds);  // D10 (+)= D11 D10
  __syncthreads();
  MpSmall::run(d01, ld, d11, ld, d01, ld, n0, n1, n1, true, flags, lds);  // D01 (+)= D01 D11
  __syncthreads();
  MpSmall::run(d01, ld, d10, ld, d00, ld, n0, n0, n1, true, flags, lds);  // D00 (+)= D01 D10
  __syncthreads();
}

// Phase 2 for T = 128 in 64 x 64 pieces: each panel tile (r, J) or (I, r)
// is split into four 64 x 64 output quarters, one workgroup each (K = 128,
// 4 x 4 cells per lane), so the phase runs 4 (2 nb - 2) short workgroups
// instead of 2 nb - 2 long ones (n = 8192: 19 us per round against 40).
// A quarter reads the other row (column) half of its own panel tile while a
// sibling writes it; every value it can see there is the length of a path
// through pivot block r no shorter than the finished one, so the minimum it
// forms is the same either way.  (Staging the whole 128-deep operands in
// LDS at once, 67 KB per workgroup, measured 21 us.)
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void fw_panel128_kernel(
    int32_t *d, int n, int64_t ld, int r, int nb, int flags) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MpSmall::LDS_WORDS];
  const int base = r * 128;
  const int kb = imin(128, n - base);
  const int32_t *diag = d + (int64_t)base * ld + base;
  const int q = blockIdx.x & 3, hr = q >> 1, hc = q & 1;
  int bid = blockIdx.x >> 2;
  if (bid < nb - 1) {
    const int J = bid < r ? bid : bid + 1;
    const int rows = kb - 64 * hr, cols = imin(128, n - J * 128) - 64 * hc;
    if (rows <= 0 || cols <= 0) return;
    int32_t *c = d + (int64_t)base * ld + (int64_t)J * 128;
    MpSmall::run(diag + (int64_t)(64 * hr) * ld, ld, c + 64 * hc, ld,
                 c + (int64_t)(64 * hr) * ld + 64 * hc, ld, imin(64, rows), imin(64, cols), kb,
                 true, flags, lds);
  } else {
    bid -= nb - 1;
    const int I = bid < r ? bid : bid + 1;
    const int rows = imin(128, n - I * 128) - 64 * hr, cols = kb - 64 * hc;
    if (rows <= 0 || cols <= 0) return;
    int32_t *c = d + (int64_t)I * 128 * ld + base;
    MpSmall::run(c + (int64_t)(64 * hr) * ld, ld, diag + 64 * hc, ld,
                 c + (int64_t)(64 * hr) * ld + 64 * hc, ld, imin(64, rows), imin(64, cols), kb,
                 true, flags, lds);
  }
}

// Phase 3 over a set of tiles, selected by `mode` (p = r + 1):
//   0: every tile outside row r and column r, (nb-1)^2 of them;
//   1: the look-ahead tiles, row p and column p minus row/column r (2nb-3);
//   2: the rest, outside rows/columns r and p, (nb-2)^2.
// Modes 1 and 2 together equal mode 0; splitting them lets the next round's
// diagonal and panels (which only touch row/column p) start while mode 2
// runs.
//
// With `fuse` (mode 0, r + 1 < nb) workgroup 0 takes tile (r+1, r+1) --
// swapping with the workgroup that had it -- and after its update closes
// it (in LDS for T = 64, by close_block128 for T = 128): the next round's
// diagonal phase, done while the other tiles of this phase run, so that
// round launches only its panels.
template <int TM>
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void fw_rest_kernel(
    int32_t *d, int n, int64_t ld, int r, int nb, int flags, int mode, int fuse) {
  using Tile = MpTile<TM, TM>;
  constexpr int T = Tile::BM;
  constexpr int LDS_WORDS = BLK_LDS_WORDS > Tile::LDS_WORDS ? BLK_LDS_WORDS : Tile::LDS_WORDS;
  __shared__ __attribute__((aligned(16))) int32_t lds[LDS_WORDS];
  int I, J;
  if (mode == 0) {
    const int m = nb - 1;
    int t = xcd_remap(blockIdx.x, m * m);
    if (fuse) {
      const int tp = r * m + r;  // tile (r+1, r+1) in the compressed index
      if (blockIdx.x == 0) t = tp;
      else if (t == tp) t = xcd_remap(0, m * m);
    }
    I = t / m;
    J = t - I * m;
    I += (I >= r);
    J += (J >= r);
  } else if (mode == 1) {
    const int p = r + 1;
    const int t = blockIdx.x;
    if (t < nb - 1) {
      I = p;
      J = t + (t >= r);
    } else {
      const int u = t - (nb - 1);
      I = u < r ? u : u + 2;
      J = p;
    }
  } else {
    const int m = nb - 2;
    const int t = xcd_remap(blockIdx.x, m * m);
    I = t / m;
    J = t - I * m;
    I = I < r ? I : I + 2;
    J = J < r ? J : J + 2;
  }
  const int base = r * T;
  const int kb = imin(T, n - base);
  int32_t *c = d + (int64_t)I * T * ld + (int64_t)J * T;
  const int32_t *a = d + (int64_t)I * T * ld + base;
  const int32_t *b = d + (int64_t)base * ld + (int64_t)J * T;
  Tile::run(a, ld, b, ld, c, ld, imin(T, n - I * T), imin(T, n - J * T), kb, true, flags, lds);
  if (fuse && blockIdx.x == 0) {
    __syncthreads();  // the tile's writes above come before the closure reads it
    if (TM == 4) close_block64(c, ld, imin(T, n - I * T), lds);
    else close_block128(c, ld, imin(T, n - I * T), flags, lds);
  }
}

}  // namespace fapk

namespace fapi {

int vec_flags(const void *A, int64_t lda, const void *B, int64_t ldb, const void *C,
              int64_t ldc);

namespace {

template <int TM>
struct Fw {
  static constexpr int T = 16 * TM;

  static int diag_panel(int32_t *d, int32_t n, int64_t ld, int r, int nb, int flags,
                        hipStream_t s, bool diag = true) {
    if (diag) {
      hipLaunchKernelGGL(fapk::fw_diag_kernel<T>, dim3(1), dim3(fapk::FW_DIAG_THREADS), 0, s,
                         d, n, ld, r);
      FAP_CHECK_LAUNCH();
    }
    if (nb > 1) {
      if constexpr (TM == 8)
        hipLaunchKernelGGL(fapk::fw_panel128_kernel, dim3(8 * (nb - 1)), dim3(fapk::MP_THREADS),
                           0, s, d, n, ld, r, nb, flags);
      else
        hipLaunchKernelGGL(fapk::fw_panel_kernel<TM>, dim3(2 * (nb - 1)),
                           dim3(fapk::MP_THREADS), 0, s, d, n, ld, r, nb, flags);
      FAP_CHECK_LAUNCH();
    }
    return FAP_OK;
  }

  static int rest(int32_t *d, int32_t n, int64_t ld, int r, int nb, int flags, int mode,
                  hipStream_t s, int fuse = 0) {
    const int tiles =
        mode == 0 ? (nb - 1) * (nb - 1) : mode == 1 ? 2 * nb - 3 : (nb - 2) * (nb - 2);
    if (tiles <= 0) return FAP_OK;
    hipLaunchKernelGGL(fapk::fw_rest_kernel<TM>, dim3(tiles), dim3(fapk::MP_THREADS), 0, s, d,
                       n, ld, r, nb, flags, mode, fuse);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }

  // With look-ahead, round r's phase 3 is split: the tiles of row/column
  // r+1 first (mode 1), then, on the auxiliary stream, round r+1's diagonal
  // and panels run while the main stream finishes round r's other tiles
  // (mode 2).  All work is ordered before later work on `s` on return.
  static int run(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, bool lookahead,
                 hipStream_t s) {
    const int nb = (n + T - 1) / T;
    const int flags = vec_flags(d, ld, d, ld, d, ld);
    FAP_TRY(diag_panel(d, n, ld, 0, nb, flags, s));
    const bool overlap = lookahead && ctx && ctx->aux_stream && nb >= 3;
    // fuse the next diagonal closure into phase 3 when the phase is long
    // enough to hide it: at least 8 tiles per CU for T = 64 (a 256-thread
    // 64-closure), 4 for T = 128 (the 128-closure's seven steps)
    const int64_t cus = ctx ? ctx->num_cus : 256;
    const bool fuse_ok = (int64_t)(nb - 1) * (nb - 1) >= cus * (TM == 4 ? 8 : 4);
    for (int r = 0; r < nb; r++) {
      if (r + 1 < nb && overlap) {
        FAP_TRY(rest(d, n, ld, r, nb, flags, 1, s));
        FAP_HIP(hipEventRecord(ctx->ev_fw_la, s));
        FAP_HIP(hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fw_la, 0));
        FAP_TRY(diag_panel(d, n, ld, r + 1, nb, flags, ctx->aux_stream));
        FAP_HIP(hipEventRecord(ctx->ev_fw_panel, ctx->aux_stream));
        FAP_TRY(rest(d, n, ld, r, nb, flags, 2, s));
        FAP_HIP(hipStreamWaitEvent(s, ctx->ev_fw_panel, 0));
      } else {
        // round r's phase 3 also closes the next diagonal block when the
        // phase is long enough to hide it (below that the 1024-thread
        // diagonal kernel is faster: T = 64, n = 1024 took 0.56 ms unfused,
        // 0.64 ms fused)
        const int fuse = (fuse_ok && r + 1 < nb) ? 1 : 0;
        FAP_TRY(rest(d, n, ld, r, nb, flags, 0, s, fuse));
        if (r + 1 < nb) FAP_TRY(diag_panel(d, n, ld, r + 1, nb, flags, s, !fuse));
      }
    }
    return FAP_OK;
  }
};

}  // namespace

// Block size and look-ahead: the context options if set, else the tune
// cache, else 128 when phase 3 has at least 8 tiles of 128 per CU and 64
// below, without look-ahead (fw_settings).  Measured on MI355X (closure
// ms, T = 64 / 128): n = 8192 20.7 / 19.0, 6144 9.49 / 8.86, 5120 5.99 /
// 6.54, 4096 3.59 / 4.87, 2048 1.40 / 1.83 -- the 128 x 128 tile with 8 x 8
// cells per lane halves the LDS bytes per update of the 64 x 64 tile's
// 4 x 4 cells, and with the next diagonal closed inside phase 3 and the
// panels split in quarters, the serial phases of T = 128 no longer cost
// more than that saves once the grid is large.
int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s) {
  if (n <= 0) return FAP_OK;
  int tile;
  bool lookahead;
  fw_settings(ctx, n, &tile, &lookahead);
  if (tile == 128) return Fw<8>::run(ctx, d, n, ld, lookahead, s);
  return Fw<4>::run(ctx, d, n, ld, lookahead, s);
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_fw_blocked(fap_ctx *ctx, int32_t *d_mat, int32_t n, int64_t ld,
                              void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (n < 0) return fail(FAP_E_ARG, "fw: negative n");
  if (n == 0) return FAP_OK;
  if (!d_mat) return fail(FAP_E_ARG, "fw: null matrix");
  if (ld < n) return fail(FAP_E_ARG, "fw: ld < n");
  if ((int64_t)(n + 63) / 64 > 46340) return fail(FAP_E_ARG, "fw: n too large");
  return launch_fw(ctx, d_mat, n, ld, pick_stream(ctx, stream));
}
