// floyd.hip -- blocked Floyd-Warshall closure on the VALU, in place.
//
// Replaces the reference's stage-2 operator floyd_path (fap/kernel/floyd.h:
// 150-158; CPU semantics floyd_CPU_path :38-49: k-i-j order, strict relax).
// The closure of a non-negative matrix is unique, so the blocked order gives
// the same distances as the k-i-j loop.  With block size T = 128 and
// nb = ceil(n / T) blocks, round r (pivot block r) runs:
//   1. diagonal tile (r, r): FW over its kb <= T pivots in one workgroup,
//      cells in registers, pivot row/column exchanged through LDS;
//   2. row panel (r, J) = min(C, D_rr (x) C) and column panel
//      (I, r) = min(C, C (x) D_rr) for all I, J != r: one launch;
//   3. every other tile (I, J) = min(C, D_Ir (x) D_rJ): one launch.
// Phases 2 and 3 reuse the register-blocked min-plus tile (minplus_tile.h).
#include "common.h"
#include "minplus_tile.h"

namespace fapk {

constexpr int FW_T = 128;
constexpr int FW_DIAG_THREADS = 1024;

// Phase 1.  Thread (ty, tx) = (t >> 5, t & 31) owns cells (ty + 32a, tx + 32b),
// a, b < 4.  Before step k the owners of row k / column k publish them into
// rowbuf / colbuf (double-buffered by k parity), one barrier per step.  Row k
// and column k do not change during step k because the diagonal is >= 0.
__global__ __launch_bounds__(FW_DIAG_THREADS) void fw_diag_kernel(int32_t *d, int n,
                                                                  int64_t ld, int r) {
  __shared__ int32_t rowbuf[2][FW_T];
  __shared__ int32_t colbuf[2][FW_T];
  const int t = threadIdx.x;
  const int ty = t >> 5, tx = t & 31;
  const int base = r * FW_T;
  const int kb = imin(FW_T, n - base);
  int32_t *blk = d + (int64_t)base * ld + base;
  int cell[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      cell[a][b] = (i < kb && j < kb) ? blk[(int64_t)i * ld + j] : kInf;
    }
  // publish pivot 0 (row 0 is owned by ty == 0, a == 0; column 0 by tx == 0, b == 0)
  if (ty == 0) {
#pragma unroll
    for (int b = 0; b < 4; b++) rowbuf[0][tx + 32 * b] = cell[0][b];
  }
  if (tx == 0) {
#pragma unroll
    for (int a = 0; a < 4; a++) colbuf[0][ty + 32 * a] = cell[a][0];
  }
  __syncthreads();
  for (int k = 0; k < kb; k++) {
    const int p = k & 1;
    int pr[4], pc[4];
#pragma unroll
    for (int b = 0; b < 4; b++) pr[b] = rowbuf[p][tx + 32 * b];
#pragma unroll
    for (int a = 0; a < 4; a++) pc[a] = colbuf[p][ty + 32 * a];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) cell[a][b] = imin(cell[a][b], pc[a] + pr[b]);
    const int nx = k + 1;
    if (nx < kb) {
      const int q = nx & 1;
      if ((nx & 31) == ty) {
        const int asel = nx >> 5;
#pragma unroll
        for (int a = 0; a < 4; a++)
          if (a == asel) {
#pragma unroll
            for (int b = 0; b < 4; b++) rowbuf[q][tx + 32 * b] = cell[a][b];
          }
      }
      if ((nx & 31) == tx) {
        const int bsel = nx >> 5;
#pragma unroll
        for (int b = 0; b < 4; b++)
          if (b == bsel) {
#pragma unroll
            for (int a = 0; a < 4; a++) colbuf[q][ty + 32 * a] = cell[a][b];
          }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      if (i < kb && j < kb) blk[(int64_t)i * ld + j] = cell[a][b];
    }
}

// Phase 2: workgroups [0, nb-1) take row-panel tiles (r, J), the next nb-1
// take column-panel tiles (I, r).  Each tile is owned by one workgroup and
// aliases one operand; minplus_tile reads all operands before writing C.
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void fw_panel_kernel(int32_t *d, int n, int64_t ld,
                                                              int r, int nb, int flags) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MP_LDS_WORDS];
  const int base = r * FW_T;
  const int kb = imin(FW_T, n - base);
  const int32_t *diag = d + (int64_t)base * ld + base;
  int bid = blockIdx.x;
  if (bid < nb - 1) {
    const int J = bid < r ? bid : bid + 1;
    int32_t *c = d + (int64_t)base * ld + (int64_t)J * FW_T;
    minplus_tile(diag, ld, c, ld, c, ld, kb, imin(FW_T, n - J * FW_T), kb, true, flags, lds);
  } else {
    bid -= nb - 1;
    const int I = bid < r ? bid : bid + 1;
    int32_t *c = d + (int64_t)I * FW_T * ld + base;
    minplus_tile(c, ld, diag, ld, c, ld, imin(FW_T, n - I * FW_T), kb, kb, true, flags, lds);
  }
}

// Phase 3: the (nb-1)^2 tiles outside row r and column r.
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void fw_rest_kernel(int32_t *d, int n, int64_t ld,
                                                             int r, int nb, int flags) {
  __shared__ __attribute__((aligned(16))) int32_t lds[MP_LDS_WORDS];
  const int m = nb - 1;
  const int t = xcd_remap(blockIdx.x, m * m);
  int I = t / m, J = t - (t / m) * m;
  I += (I >= r);
  J += (J >= r);
  const int base = r * FW_T;
  const int kb = imin(FW_T, n - base);
  int32_t *c = d + (int64_t)I * FW_T * ld + (int64_t)J * FW_T;
  const int32_t *a = d + (int64_t)I * FW_T * ld + base;
  const int32_t *b = d + (int64_t)base * ld + (int64_t)J * FW_T;
  minplus_tile(a, ld, b, ld, c, ld, imin(FW_T, n - I * FW_T), imin(FW_T, n - J * FW_T), kb,
               true, flags, lds);
}

}  // namespace fapk

namespace fapi {

int vec_flags(const void *A, int64_t lda, const void *B, int64_t ldb, const void *C,
              int64_t ldc);

int launch_fw(int32_t *d, int32_t n, int64_t ld, hipStream_t s) {
  if (n <= 0) return FAP_OK;
  const int nb = (n + fapk::FW_T - 1) / fapk::FW_T;
  const int flags = vec_flags(d, ld, d, ld, d, ld);
  for (int r = 0; r < nb; r++) {
    hipLaunchKernelGGL(fapk::fw_diag_kernel, dim3(1), dim3(fapk::FW_DIAG_THREADS), 0, s, d, n,
                       ld, r);
    FAP_CHECK_LAUNCH();
    if (nb > 1) {
      hipLaunchKernelGGL(fapk::fw_panel_kernel, dim3(2 * (nb - 1)), dim3(fapk::MP_THREADS), 0,
                         s, d, n, ld, r, nb, flags);
      FAP_CHECK_LAUNCH();
      hipLaunchKernelGGL(fapk::fw_rest_kernel, dim3((nb - 1) * (nb - 1)),
                         dim3(fapk::MP_THREADS), 0, s, d, n, ld, r, nb, flags);
      FAP_CHECK_LAUNCH();
    }
  }
  return FAP_OK;
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
  if ((int64_t)(n + fapk::FW_T - 1) / fapk::FW_T > 46340)
    return fail(FAP_E_ARG, "fw: n too large");
  return launch_fw(d_mat, n, ld, pick_stream(ctx, stream));
}
