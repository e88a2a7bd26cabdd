// path_ops.hip -- path matrices of the min-plus and Floyd-Warshall operators.
//
// The reference's operators carry a path matrix beside the distances:
//   min-plus  (fap/kernel/minplus.h:51-65): res[i][j] starts at INF and
//             path[i][j] at -1; for k ascending, a strictly smaller
//             mat1[i][k] + mat2[k][j] replaces res and sets
//             path[i][j] = mat2_path[k][j].  The path is therefore
//             mat2_path[k*][j] for the FIRST k attaining the minimum, -1 when
//             the result stays INF -- a deterministic function of the inputs,
//             reproduced here exactly.
//   FW        (fap/kernel/floyd.h:38-49): a strict improvement through k sets
//             path[i][j] = path[k][j].  Which k wins among equal paths
//             depends on the relaxation order (the reference's CPU, CUDA and
//             HIP backends differ), so what is reproduced is the contract:
//             a cell the closure did not improve keeps its input label; an
//             improved cell takes path_in[m][j] for a last hop m -> j of a
//             shortest i -> j path in the input matrix (d[i][m] +
//             d_in[m][j] == d[i][j]), chosen so that following the hops
//             m ends at i.
//
// FW hops are read off the closed matrix afterwards, like the predecessor
// rows of path.hip: one pass takes, per improved cell, the smallest m with a
// positive tight hop (d_in[m][j] > 0, so d[i][m] < d[i][j]); cells whose only
// tight hops have weight 0 (equal-distance plateaus) are resolved in further
// passes, each taking a zero-weight tight hop from a cell resolved in an
// earlier pass, so hop chains end at resolved cells and, through them, at i.
// Every choice is the smallest eligible m, so the labels are a function of
// the inputs alone (the same on every run).
#include <algorithm>

#include "common.h"

namespace fapk {

constexpr int kPT = 64;   // output tile of the min-plus path kernel
constexpr int kPK = 16;   // k chunk staged in LDS

// C (M x N) = min(INF, A (M x K) (x) B (K x N)); Cp[i][j] = Bp[k*][j] for
// the first minimising k, -1 when C[i][j] is INF.  256 threads, 64 x 64
// outputs per workgroup, 4 x 4 per lane (rows ty + 16 r, columns tx + 16 c).
// Out-of-range operands read INF; INF + INF = 2^31 - 2 does not overflow and
// never beats the INF start.
__global__ __launch_bounds__(256) void minplus_path_kernel(
    const int32_t *__restrict__ A, int64_t lda, const int32_t *__restrict__ B, int64_t ldb,
    const int32_t *__restrict__ Bp, int64_t ldbp, int32_t *__restrict__ C, int64_t ldc,
    int32_t *__restrict__ Cp, int64_t ldcp, int32_t M, int32_t N, int32_t K) {
  __shared__ int32_t As[kPK][kPT + 1];
  __shared__ int32_t Bs[kPK][kPT];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int64_t m0 = (int64_t)blockIdx.y * kPT, n0 = (int64_t)blockIdx.x * kPT;
  int32_t best[4][4], bk[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      best[r][c] = FAP_INF;
      bk[r][c] = -1;
    }
  for (int k0 = 0; k0 < K; k0 += kPK) {
    for (int e = tid; e < kPT * kPK; e += 256) {
      const int r = e / kPK, kk = e % kPK;
      const int64_t gi = m0 + r;
      const int gk = k0 + kk;
      As[kk][r] = (gi < M && gk < K) ? A[gi * lda + gk] : FAP_INF;
    }
    for (int e = tid; e < kPT * kPK; e += 256) {
      const int kk = e / kPT, c = e % kPT;
      const int gk = k0 + kk;
      const int64_t gj = n0 + c;
      Bs[kk][c] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : FAP_INF;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kPK; kk++) {
      int32_t a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; r++) a[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int32_t s = a[r] + b[c];
          if (s < best[r][c]) {
            best[r][c] = s;
            bk[r][c] = k0 + kk;
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int64_t i = m0 + ty + 16 * r;
    if (i >= M) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int64_t j = n0 + tx + 16 * c;
      if (j >= N) continue;
      C[i * ldc + j] = best[r][c];
      if (Cp) Cp[i * ldcp + j] = best[r][c] < FAP_INF ? Bp[(int64_t)bk[r][c] * ldbp + j] : -1;
    }
  }
}

// Cell states of the FW hop passes: open, resolved by the first pass (1), or
// resolved by plateau pass q (2 + q).
constexpr int32_t kHopOpen = 0;  // improved cell without a hop yet
constexpr int32_t kHopDone = 1;

// First pass over the closed n x n matrix d (input copy d0, labels p0):
// unimproved cells keep their label; improved cells take the smallest m with
// a positive tight hop.  Counts the cells left open.
__global__ void fw_hop_kernel(const int32_t *__restrict__ d0, int64_t ld0,
                              const int32_t *__restrict__ d, int64_t ld,
                              const int32_t *__restrict__ p0, int64_t ldp0,
                              int32_t *__restrict__ p, int64_t ldp, int32_t *__restrict__ state,
                              int32_t n, unsigned int *__restrict__ open) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  unsigned int left = 0;
  for (int i = blockIdx.y; i < n; i += gridDim.y) {
    const int32_t *di = d + (int64_t)i * ld;
    const int32_t dij = di[j];
    int32_t lab, st = kHopDone;
    if (dij == d0[(int64_t)i * ld0 + j]) {
      lab = p0[(int64_t)i * ldp0 + j];
    } else {
      int m = -1;
      for (int mm = 0; mm < n; mm++) {
        const int32_t w = d0[(int64_t)mm * ld0 + j];
        if (w > 0 && w < FAP_INF && di[mm] + w == dij) {
          m = mm;
          break;
        }
      }
      if (m >= 0) {
        lab = p0[(int64_t)m * ldp0 + j];
      } else {
        lab = -1;
        st = kHopOpen;
        left++;
      }
    }
    p[(int64_t)i * ldp + j] = lab;
    state[(int64_t)i * n + j] = st;
  }
  if (left) atomicAdd(open, left);
}

// One plateau pass (stamp = 2 + pass number): an open cell takes the
// smallest m != j with a zero-weight tight hop whose own cell (i, m) was
// resolved BEFORE this pass (1 <= state < stamp), and is stamped.  Cells
// resolved during the pass are not used until the next one, so the choice
// does not depend on the order in which workgroups run.  Counts the cells
// resolved.
__global__ void fw_hop_plateau_kernel(const int32_t *__restrict__ d0, int64_t ld0,
                                      const int32_t *__restrict__ d, int64_t ld,
                                      const int32_t *__restrict__ p0, int64_t ldp0,
                                      int32_t *__restrict__ p, int64_t ldp, int32_t *state,
                                      int32_t n, int32_t stamp,
                                      unsigned int *__restrict__ resolved) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  unsigned int got = 0;
  for (int i = blockIdx.y; i < n; i += gridDim.y) {
    int32_t *si = state + (int64_t)i * n;
    if (__hip_atomic_load(si + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != kHopOpen)
      continue;
    const int32_t *di = d + (int64_t)i * ld;
    const int32_t dij = di[j];
    for (int mm = 0; mm < n; mm++) {
      if (mm == j || d0[(int64_t)mm * ld0 + j] != 0 || di[mm] != dij) continue;
      const int32_t sm = __hip_atomic_load(si + mm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (sm < kHopDone || sm >= stamp) continue;
      p[(int64_t)i * ldp + j] = p0[(int64_t)mm * ldp0 + j];
      __hip_atomic_store(si + j, stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      got++;
      break;
    }
  }
  if (got) atomicAdd(resolved, got);
}

}  // namespace fapk

namespace fapi {

int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);

int run_minplus_path(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb,
                     const int32_t *Bp, int64_t ldbp, int32_t *C, int64_t ldc, int32_t *Cp,
                     int64_t ldcp, int32_t M, int32_t N, int32_t K, hipStream_t s) {
  if (M <= 0 || N <= 0) return FAP_OK;
  if ((int64_t)(M + fapk::kPT - 1) / fapk::kPT > 65535)
    return fail(FAP_E_ARG, "min-plus path: M too large for one launch");
  dim3 grd((unsigned)((N + fapk::kPT - 1) / fapk::kPT), (unsigned)((M + fapk::kPT - 1) / fapk::kPT));
  hipLaunchKernelGGL(fapk::minplus_path_kernel, grd, dim3(256), 0, s, A, lda, B, ldb, Bp, ldbp,
                     C, ldc, Cp, ldcp, M, N, K);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// Closes d (n x n, stride ld) in place and writes the path labels: p holds
// the input labels on entry and the output labels on return (stride ldp).
// Workspace: a copy of the input matrix, of the input labels and the cell
// states, 12 n^2 bytes, in ctx->ws_misc.  Synchronizes the stream (the
// plateau passes are driven from the host).
int run_fw_path(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, int32_t *p, int64_t ldp,
                hipStream_t s) {
  if (n <= 0) return FAP_OK;
  const int64_t n2 = (int64_t)n * n;
  FAP_TRY(ensure_ws(ctx, ctx->ws_misc, 256 + 3 * (size_t)n2 * sizeof(int32_t)));
  unsigned int *cnt = (unsigned int *)ctx->ws_misc.ptr;
  int32_t *d0 = (int32_t *)((char *)ctx->ws_misc.ptr + 256);
  int32_t *p0 = d0 + n2;
  int32_t *state = p0 + n2;
  FAP_HIP(hipMemcpy2DAsync(d0, sizeof(int32_t) * n, d, sizeof(int32_t) * ld,
                           sizeof(int32_t) * n, n, hipMemcpyDeviceToDevice, s));
  FAP_HIP(hipMemcpy2DAsync(p0, sizeof(int32_t) * n, p, sizeof(int32_t) * ldp,
                           sizeof(int32_t) * n, n, hipMemcpyDeviceToDevice, s));
  FAP_TRY(launch_fw(ctx, d, n, ld, s));
  dim3 blk(256), grd((n + 255) / 256, std::min(n, 1024));
  FAP_TRY(launch_fill_u32(cnt, 1, 0u, s));
  hipLaunchKernelGGL(fapk::fw_hop_kernel, grd, blk, 0, s, d0, (int64_t)n, d, ld, p0, (int64_t)n,
                     p, ldp, state, n, cnt);
  FAP_CHECK_LAUNCH();
  unsigned int h = 0;
  FAP_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  int64_t left = h;
  // every pass resolves at least one open cell of each row that has one (the
  // open cell with the fewest hops on a shortest path has a zero-weight tight
  // hop from a cell resolved before the pass), so n passes bound the loop
  for (int pass = 0; left > 0; pass++) {
    if (pass >= n) return fail(FAP_E_STATE, "FW path plateaus did not resolve");
    FAP_TRY(launch_fill_u32(cnt, 1, 0u, s));
    hipLaunchKernelGGL(fapk::fw_hop_plateau_kernel, grd, blk, 0, s, d0, (int64_t)n, d, ld, p0,
                       (int64_t)n, p, ldp, state, n, (int32_t)(2 + pass), cnt);
    FAP_CHECK_LAUNCH();
    FAP_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    FAP_HIP(hipStreamSynchronize(s));
    if (h == 0) return fail(FAP_E_STATE, "FW path: an improved cell has no tight hop");
    left -= h;
  }
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_minplus_path(fap_ctx *ctx, const int32_t *A, int64_t lda, const int32_t *B,
                                int64_t ldb, const int32_t *B_path, int64_t ldbp, int32_t *C,
                                int64_t ldc, int32_t *C_path, int64_t ldcp, int32_t M, int32_t N,
                                int32_t K, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "null ctx");
  if (M < 0 || N < 0 || K < 0) return fail(FAP_E_ARG, "min-plus path: negative size");
  if (M == 0 || N == 0) return FAP_OK;
  if (!C || (K > 0 && (!A || !B)) || (C_path && K > 0 && !B_path))
    return fail(FAP_E_ARG, "min-plus path: null operand");
  if (ldc < N || (K > 0 && (lda < K || ldb < N)) || (C_path && ldcp < N) ||
      (C_path && K > 0 && ldbp < N))
    return fail(FAP_E_ARG, "min-plus path: leading dimension below the row length");
  (void)hipSetDevice(ctx->device);
  return run_minplus_path(A, lda, B, ldb, B_path, ldbp, C, ldc, C_path, ldcp, M, N, K,
                          pick_stream(ctx, stream));
}

extern "C" int fap_fw_blocked_path(fap_ctx *ctx, int32_t *d_mat, int32_t n, int64_t ld,
                                   int32_t *d_path, int64_t ldp, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "null ctx");
  if (n < 0) return fail(FAP_E_ARG, "FW path: negative size");
  if (n == 0) return FAP_OK;
  if (!d_mat || !d_path || ld < n || ldp < n)
    return fail(FAP_E_ARG, "FW path: bad pointers or strides");
  (void)hipSetDevice(ctx->device);
  return run_fw_path(ctx, d_mat, n, ld, d_path, ldp, pick_stream(ctx, stream));
}
