// submat.hip -- part glue on the device and the per-part pipeline.
//
// gather  <- MysubMatBuild_path  fap/utils/util_Myalgorithm.h:30-73
// (view)  <- MyMat1Build         fap/utils/util_Myalgorithm.h:88-95: the
//            inner x bdy operand is read in place as rows n_bdy.. of the part
//            matrix with its own stride, no copy
// scatter <- MysubMatDecode_path fap/utils/util_Myalgorithm.h:75-86
// solve   <- fapGraph::run       fap/fap.cpp:92-181
#include <algorithm>

#include "common.h"

namespace fapk {

constexpr int GS_BX = 64, GS_BY = 4;

// Rows < n_bdy: SSSP distances gathered through st2ed; other rows INF;
// diagonal 0.  (The reference also writes the intra-part arcs of boundary
// rows and then overwrites those rows with the SSSP distances, :36-65.)
__global__ void gather_fill_kernel(const int32_t *__restrict__ st2ed, int start, int n_sub,
                                   int n_bdy, const int32_t *__restrict__ rowblock,
                                   int64_t ld_rb, int32_t *__restrict__ sub, int64_t ld_sub) {
  const int j = blockIdx.x * GS_BX + threadIdx.x;
  if (j >= n_sub) return;
  const int vj = st2ed[start + j];
  for (int i = blockIdx.y * GS_BY + threadIdx.y; i < n_sub; i += gridDim.y * GS_BY) {
    int v;
    if (i == j) {
      v = 0;
    } else if (i < n_bdy) {
      v = rowblock[(int64_t)i * ld_rb + vj];
    } else {
      v = 0x3FFFFFFF;
    }
    sub[(int64_t)i * ld_sub + j] = v;
  }
}

// Inner rows: the part's own arcs in CSR order, one thread per row so that
// the last duplicate arc wins (:40-52); diagonal forced to 0 afterwards
// (:67-72) so a self loop cannot override it.
__global__ void gather_arcs_kernel(const int32_t *__restrict__ row_offset,
                                   const int2 *__restrict__ edges,
                                   const int32_t *__restrict__ graph_id,
                                   const int32_t *__restrict__ st2ed,
                                   const int32_t *__restrict__ ed2st, int start, int n_sub,
                                   int n_bdy, int32_t *__restrict__ sub, int64_t ld_sub) {
  const int i = n_bdy + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_sub) return;
  const int u = st2ed[start + i];
  const int gu = graph_id[u];
  int32_t *row = sub + (int64_t)i * ld_sub;
  for (int e = row_offset[u]; e < row_offset[u + 1]; e++) {
    const int2 ed = edges[e];
    if (graph_id[ed.x] != gu) continue;
    const int j = ed2st[ed.x] - start;
    if (j >= 0 && j < n_sub) row[j] = ed.y;
  }
  row[i] = 0;
}

__global__ void scatter_kernel(const int32_t *__restrict__ st2ed, int start, int rows, int n_sub,
                               const int32_t *__restrict__ sub, int64_t ld_sub,
                               int32_t *__restrict__ rowblock, int64_t ld_rb) {
  const int j = blockIdx.x * GS_BX + threadIdx.x;
  if (j >= n_sub) return;
  const int vj = st2ed[start + j];
  for (int i = blockIdx.y * GS_BY + threadIdx.y; i < rows; i += gridDim.y * GS_BY)
    rowblock[(int64_t)i * ld_rb + vj] = sub[(int64_t)i * ld_sub + j];
}

}  // namespace fapk

namespace fapi {

int run_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
             int32_t *d_out, int64_t ld_out, hipStream_t s);
int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);
int launch_minplus(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C,
                   int64_t ldc, int32_t M, int32_t N, int32_t K, int accumulate,
                   hipStream_t s);

static int check_part(const fap_graph_dev *g, int32_t start, int32_t n_sub, int32_t n_bdy) {
  if (!g->graph_id || !g->st2ed || !g->ed2st)
    return fail(FAP_E_STATE, "graph has no partition (fap_graph_set_partition)");
  if (start < 0 || n_sub < 0 || n_bdy < 0 || n_bdy > n_sub || (int64_t)start + n_sub > g->V)
    return fail(FAP_E_ARG, "part range outside the vertex permutation");
  return FAP_OK;
}

int launch_gather(const fap_graph_dev *g, int32_t start, int32_t n_sub, int32_t n_bdy,
                  const int32_t *rowblock, int64_t ld_rb, int32_t *sub, int64_t ld_sub,
                  hipStream_t s) {
  if (n_sub == 0) return FAP_OK;
  dim3 blk(fapk::GS_BX, fapk::GS_BY);
  dim3 grd((n_sub + fapk::GS_BX - 1) / fapk::GS_BX,
           std::min((n_sub + fapk::GS_BY - 1) / fapk::GS_BY, 1024));
  hipLaunchKernelGGL(fapk::gather_fill_kernel, grd, blk, 0, s, g->st2ed, start, n_sub, n_bdy,
                     rowblock, ld_rb, sub, ld_sub);
  FAP_CHECK_LAUNCH();
  const int inner = n_sub - n_bdy;
  if (inner > 0) {
    hipLaunchKernelGGL(fapk::gather_arcs_kernel, dim3((inner + 255) / 256), dim3(256), 0, s,
                       g->row_offset, g->edges, g->graph_id, g->st2ed, g->ed2st, start, n_sub,
                       n_bdy, sub, ld_sub);
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

int launch_scatter(const fap_graph_dev *g, int32_t start, int32_t rows, int32_t n_sub,
                   const int32_t *sub, int64_t ld_sub, int32_t *rowblock, int64_t ld_rb,
                   hipStream_t s) {
  if (rows == 0 || n_sub == 0) return FAP_OK;
  dim3 blk(fapk::GS_BX, fapk::GS_BY);
  dim3 grd((n_sub + fapk::GS_BX - 1) / fapk::GS_BX, std::min((rows + fapk::GS_BY - 1) / fapk::GS_BY, 1024));
  hipLaunchKernelGGL(fapk::scatter_kernel, grd, blk, 0, s, g->st2ed, start, rows, n_sub, sub,
                     ld_sub, rowblock, ld_rb);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int run_subgraph_solve(fap_ctx *ctx, const fap_graph_dev *g, int32_t start, int32_t n_sub,
                       int32_t n_bdy, int32_t *rowblock, int64_t ld_rb, hipStream_t s) {
  if (n_sub == 0) return FAP_OK;
  const int64_t ld_sub = ((int64_t)n_sub + 3) & ~(int64_t)3;
  FAP_TRY(ensure_ws(ctx, ctx->ws_sub, (size_t)ld_sub * n_sub * sizeof(int32_t)));
  int32_t *sub = (int32_t *)ctx->ws_sub.ptr;
  // stage 1: rows 0..n_bdy-1 = SSSP from the boundary vertices st2ed[start..)
  FAP_TRY(run_sssp(ctx, g, g->st2ed + start, n_bdy, rowblock, ld_rb, s));
  // stage 2: dense part matrix and its closure
  FAP_TRY(launch_gather(g, start, n_sub, n_bdy, rowblock, ld_rb, sub, ld_sub, s));
  FAP_TRY(launch_fw(ctx, sub, n_sub, ld_sub, s));
  // stage 3: inner rows = (inner x bdy) (x) (bdy x V), starting from INF
  const int32_t inner = n_sub - n_bdy;
  FAP_TRY(launch_minplus(sub + (int64_t)n_bdy * ld_sub, ld_sub, rowblock, ld_rb,
                         rowblock + (int64_t)n_bdy * ld_rb, ld_rb, inner, g->V, n_bdy, 0, s));
  // the part's own columns take the closure values
  FAP_TRY(launch_scatter(g, start, n_sub, n_sub, sub, ld_sub, rowblock, ld_rb, s));
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_gather_submat(fap_ctx *ctx, const fap_graph_dev *g, int32_t start, int32_t n_sub,
                      int32_t n_bdy, const int32_t *rowblock, int64_t ld_rb, int32_t *sub,
                      int64_t ld_sub, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  FAP_TRY(check_part(g, start, n_sub, n_bdy));
  if (n_sub > 0 && (!sub || ld_sub < n_sub)) return fail(FAP_E_ARG, "gather: bad sub/ld_sub");
  if (n_bdy > 0 && (!rowblock || ld_rb < g->V))
    return fail(FAP_E_ARG, "gather: bad rowblock/ld_rb");
  return launch_gather(g, start, n_sub, n_bdy, rowblock, ld_rb, sub, ld_sub,
                       pick_stream(ctx, stream));
}

int fap_scatter_submat(fap_ctx *ctx, const fap_graph_dev *g, int32_t start, int32_t rows,
                       int32_t n_sub, const int32_t *sub, int64_t ld_sub, int32_t *rowblock,
                       int64_t ld_rb, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  FAP_TRY(check_part(g, start, n_sub, 0));
  if (rows < 0) return fail(FAP_E_ARG, "scatter: negative rows");
  if (rows > 0 && n_sub > 0 && (!sub || !rowblock || ld_sub < n_sub || ld_rb < g->V))
    return fail(FAP_E_ARG, "scatter: bad pointers or strides");
  return launch_scatter(g, start, rows, n_sub, sub, ld_sub, rowblock, ld_rb,
                        pick_stream(ctx, stream));
}

int fap_subgraph_solve(fap_ctx *ctx, const fap_graph_dev *g, int32_t start, int32_t n_sub,
                       int32_t n_bdy, int32_t *rowblock, int64_t ld_rb, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  FAP_TRY(check_part(g, start, n_sub, n_bdy));
  if (n_sub > 0 && (!rowblock || ld_rb < g->V))
    return fail(FAP_E_ARG, "subgraph_solve: bad rowblock/ld_rb");
  return run_subgraph_solve(ctx, g, start, n_sub, n_bdy, rowblock, ld_rb,
                            pick_stream(ctx, stream));
}

}  // extern "C"
