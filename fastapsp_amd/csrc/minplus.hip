// minplus.hip -- min-plus product C = min(base, A (x) B) on the VALU.
//
// Replaces the reference's stage-3 operator (min_plus_path_advanced,
// fap/kernel/minplus.h:78-89; CPU semantics :51-65: every output starts at
// INF and takes the strict minimum over k).  One 128 x 128 tile per
// workgroup, see minplus_tile.h for the register/LDS layout.
#include "common.h"
#include "minplus_tile.h"

namespace fapk {

// Tile order: workgroups of one XCD take a contiguous tile range; inside it,
// groups of 8 row tiles walk the column panels together so that a B panel
// (K x 128) is reused by the 8 tiles that need it while it is in L2.
__device__ __forceinline__ void gemm_tile_coords(int bid, int tiles_m, int tiles_n, int *tm,
                                                 int *tn) {
  const int total = tiles_m * tiles_n;
  const int t = xcd_remap(bid, total);
  const int G = 8;
  const int group = t / (G * tiles_n);
  const int first_m = group * G;
  const int gsize = imin(G, tiles_m - first_m);
  const int within = t - group * G * tiles_n;
  *tm = first_m + within % gsize;
  *tn = within / gsize;
}

__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void minplus_gemm_kernel(
    const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C, int64_t ldc,
    int M, int N, int K, int accumulate, int flags, int tiles_m, int tiles_n) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MpBig::LDS_WORDS];
  int tm, tn;
  gemm_tile_coords(blockIdx.x, tiles_m, tiles_n, &tm, &tn);
  const int m0 = tm * MP_BM, n0 = tn * MP_BN;
  MpBig::run(A + (int64_t)m0 * lda, lda, B + n0, ldb, C + (int64_t)m0 * ldc + n0, ldc,
             imin(MP_BM, M - m0), imin(MP_BN, N - n0), K, accumulate != 0, flags, lds);
}

// The last rows of a product whose row count leaves at most 64 rows past
// its last full 128-row tile: 64 x 64 tiles, one row of them, so that the
// tail computes 64 instead of 128 rows per column panel.
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void minplus_tail64_kernel(
    const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C, int64_t ldc,
    int M, int N, int K, int accumulate, int flags, int tiles_n) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MpSmall::LDS_WORDS];
  const int tn = xcd_remap(blockIdx.x, tiles_n);
  const int n0 = tn * 64;
  MpSmall::run(A, lda, B + n0, ldb, C + n0, ldc, M, imin(64, N - n0), K, accumulate != 0, flags,
               lds);
}

// At most 16 rows of C (a short product, or the rows past a product's last
// full 128-row tile): a thread per column keeps its up to 16 cells in
// registers, so B is read once, coalesced along its rows, four k at a time;
// A's rows are staged in LDS 64 k at a time and read as wave-uniform
// (broadcast) 16-byte words.  Rows >= M and k >= K read INF (INF + INF does
// not overflow).  A c3 part's 13-row tail: ~30 instead of 82 us in 64 x 64
// tiles that computed 64 rows.
constexpr int ROWS16 = 16, ROWS16_KC = 64;
__global__ __launch_bounds__(256) void minplus_rows16_kernel(const int32_t *A, int64_t lda,
                                                            const int32_t *B, int64_t ldb,
                                                            int32_t *C, int64_t ldc, int M, int N,
                                                            int K, int accumulate) {
  __shared__ __attribute__((aligned(16))) int32_t As[ROWS16 * ROWS16_KC];
  const int col = blockIdx.x * 256 + threadIdx.x;
  const bool ok = col < N;
  int acc[ROWS16];
#pragma unroll
  for (int r = 0; r < ROWS16; r++) acc[r] = kInf;
  for (int k0 = 0; k0 < K; k0 += ROWS16_KC) {
    const int kc = imin(ROWS16_KC, K - k0);
    __syncthreads();  // the previous chunk's reads of As are done
    for (int i = threadIdx.x; i < ROWS16 * ROWS16_KC; i += 256) {
      const int r = i / ROWS16_KC, k = i % ROWS16_KC;
      As[i] = (r < M && k < kc) ? A[(int64_t)r * lda + k0 + k] : kInf;
    }
    __syncthreads();
    for (int k = 0; k < kc; k += 4) {  // k + 3 < ROWS16_KC: As holds INF past kc
      int b[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        b[q] = (ok && k + q < kc) ? B[(int64_t)(k0 + k + q) * ldb + col] : kInf;
#pragma unroll
      for (int r = 0; r < ROWS16; r++) {
        const int4 a = *reinterpret_cast<const int4 *>(As + r * ROWS16_KC + k);
        acc[r] = imin(acc[r], imin(imin(a.x + b[0], a.y + b[1]), imin(a.z + b[2], a.w + b[3])));
      }
    }
  }
  if (!ok) return;
#pragma unroll
  for (int r = 0; r < ROWS16; r++) {
    if (r >= M) break;
    int32_t *c = C + (int64_t)r * ldc + col;
    *c = accumulate ? imin(acc[r], *c) : acc[r];
  }
}

}  // namespace fapk

namespace fapi {

int vec_flags(const void *A, int64_t lda, const void *B, int64_t ldb, const void *C,
              int64_t ldc) {
  int f = 0;
  if (((uintptr_t)A & 15) == 0 && (lda & 3) == 0) f |= fapk::MP_VEC_A;
  if (((uintptr_t)B & 15) == 0 && (ldb & 3) == 0) f |= fapk::MP_VEC_B;
  if (((uintptr_t)C & 15) == 0 && (ldc & 3) == 0) f |= fapk::MP_VEC_C;
  return f;
}

int launch_minplus(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C,
                   int64_t ldc, int32_t M, int32_t N, int32_t K, int accumulate,
                   hipStream_t s) {
  if (M <= 0 || N <= 0) return FAP_OK;
  const int64_t cols256 = (N + 255) / 256;
  if (M <= fapk::ROWS16) {
    hipLaunchKernelGGL(fapk::minplus_rows16_kernel, dim3((unsigned)cols256), dim3(256), 0, s, A,
                       lda, B, ldb, C, ldc, (int)M, (int)N, (int)K, accumulate);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }
  // rows past the last full 128-row tile: a 16- or 64-row tail when they fit
  const int32_t rem = M % fapk::MP_BM;
  const int32_t tail = (M > fapk::MP_BM && rem > 0 && rem <= 64) ? rem : 0;
  const int32_t Mb = M - tail;
  const int64_t tiles_m = (Mb + fapk::MP_BM - 1) / fapk::MP_BM;
  const int64_t tiles_n = (N + fapk::MP_BN - 1) / fapk::MP_BN;
  if (tiles_m * tiles_n > INT32_MAX) return fail(FAP_E_ARG, "min-plus: too many tiles");
  const int flags = vec_flags(A, lda, B, ldb, C, ldc);
  hipLaunchKernelGGL(fapk::minplus_gemm_kernel, dim3((unsigned)(tiles_m * tiles_n)),
                     dim3(fapk::MP_THREADS), 0, s, A, lda, B, ldb, C, ldc, (int)Mb, (int)N,
                     (int)K, accumulate, flags, (int)tiles_m, (int)tiles_n);
  FAP_CHECK_LAUNCH();
  if (tail > 0 && tail <= fapk::ROWS16) {
    hipLaunchKernelGGL(fapk::minplus_rows16_kernel, dim3((unsigned)cols256), dim3(256), 0, s,
                       A + (int64_t)Mb * lda, lda, B, ldb, C + (int64_t)Mb * ldc, ldc, (int)tail,
                       (int)N, (int)K, accumulate);
    FAP_CHECK_LAUNCH();
  } else if (tail > 0) {
    const int64_t tiles_n64 = (N + 63) / 64;
    const int32_t *At = A + (int64_t)Mb * lda;
    int32_t *Ct = C + (int64_t)Mb * ldc;
    hipLaunchKernelGGL(fapk::minplus_tail64_kernel, dim3((unsigned)tiles_n64),
                       dim3(fapk::MP_THREADS), 0, s, At, lda, B, ldb, Ct, ldc, (int)tail, (int)N,
                       (int)K, accumulate, vec_flags(At, lda, B, ldb, Ct, ldc),
                       (int)tiles_n64);
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_minplus(fap_ctx *ctx, const int32_t *A, int64_t lda, const int32_t *B,
                           int64_t ldb, int32_t *C, int64_t ldc, int32_t M, int32_t N,
                           int32_t K, int accumulate, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (M < 0 || N < 0 || K < 0) return fail(FAP_E_ARG, "min-plus: negative size");
  if (M == 0 || N == 0) return FAP_OK;
  if (!C || (K > 0 && (!A || !B))) return fail(FAP_E_ARG, "min-plus: null operand");
  if (ldc < N || (K > 0 && (lda < K || ldb < N)))
    return fail(FAP_E_ARG, "min-plus: leading dimension smaller than the row length");
  return launch_minplus(A, lda, B, ldb, C, ldc, M, N, K, accumulate, pick_stream(ctx, stream));
}
