// submat.hip -- part glue on the device and the per-part pipeline.
//
// gather  <- MysubMatBuild_path  fap/utils/util_Myalgorithm.h:30-73
// (view)  <- MyMat1Build         fap/utils/util_Myalgorithm.h:88-95: the
//            inner x bdy operand is read in place as rows n_bdy.. of the part
//            matrix with its own stride, no copy
// scatter <- MysubMatDecode_path fap/utils/util_Myalgorithm.h:75-86
// solve   <- fapGraph::run       fap/fap.cpp:92-181
#include <string.h>

#include <algorithm>

#include "batch_desc.h"
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
int run_sssp_rows(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
                  int32_t *d_out, int64_t ld_out, int32_t *const *d_rowptr, hipStream_t s);
int fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *mats, const int32_t *n,
                   const int64_t *ld, int use_graph, fap_fw_plan **out);
int fw_plan_run(fap_fw_plan *p, hipStream_t s);
void fw_plan_destroy(fap_fw_plan *p);
int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);
int launch_minplus(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C,
                   int64_t ldc, int32_t M, int32_t N, int32_t K, int accumulate,
                   hipStream_t s);

static int check_part(const fap_graph_dev *g, int32_t start, int32_t n_sub, int32_t n_bdy) {
  if (!g->graph_id || !g->st2ed || !g->ed2st)
    return fail(FAP_E_STATE, "graph has no partition (fap_graph_set_partition)");
  if (start < 0 || n_sub < 0 || n_bdy < 0 || n_bdy > n_sub || (int64_t)start + n_sub > g->V)
    return fail(FAP_E_ARG, "part range outside the vertex permutation");
  // a part is a maximal run of one graph id in st2ed order (fap/fap.cpp:73-78)
  if (n_sub > 0 && (int64_t)g->h_graph_id.size() == g->V && (int64_t)g->h_st2ed.size() == g->V) {
    const auto gid = [&](int32_t k) { return g->h_graph_id[g->h_st2ed[k]]; };
    const int32_t id = gid(start), end = start + n_sub;
    if (gid(end - 1) != id || (start > 0 && gid(start - 1) == id) ||
        (end < g->V && gid(end) == id))
      return fail(FAP_E_ARG, "part range [" + std::to_string(start) + ", " +
                                 std::to_string(end) + ") is not one whole part");
  }
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

// Enqueues the batched solve of `count` parts on stream s: every boundary
// vertex's row (a row pointer per source into its part's row block) through
// the boundary index when it holds every part -- the host waits on s while
// the index's iterations run -- else one SSSP launch; the part matrices
// gathered side by side in ws_sub, one batched closure, then each part's
// product and scatter.  Sources, row pointers and closure descriptors go up
// in one copy from the context's pinned staging.  ctx's workspaces stay in
// use until s passes this work; a later enqueue on the same stream is
// ordered after it.  Stage 1 waits for `sssp_after` (if given) and is
// followed by a record of `sssp_done` (if given).
int enqueue_subgraph_batch(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                           const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                           int32_t *const *rowblocks, int64_t ld_rb, hipStream_t s,
                           hipEvent_t sssp_after = nullptr, hipEvent_t sssp_done = nullptr) {
  if (count <= 0) return FAP_OK;
  if ((int64_t)g->h_st2ed.size() != g->V)
    return fail(FAP_E_STATE, "graph has no partition (fap_graph_set_partition)");
  int64_t nsrc = 0, words = 0;
  std::vector<int64_t> ld_sub(count), off(count);
  for (int32_t p = 0; p < count; p++) {
    nsrc += n_bdys[p];
    ld_sub[p] = ((int64_t)n_subs[p] + 3) & ~(int64_t)3;
    off[p] = words;
    words += ld_sub[p] * n_subs[p];
  }
  if (nsrc > INT32_MAX) return fail(FAP_E_ARG, "subgraph batch: too many sources");
  // stage 1 through the boundary index when it holds every part as given
  std::vector<int32_t> qidx(count, -1), rowoff(count, 0);
  bool use_index = ctx->opt_sssp_index && g->bidx && nsrc > 0 && (nsrc + 63) / 64 <= 65535;
  for (int32_t p = 0, r = 0; p < count && use_index; r += n_bdys[p], p++) {
    qidx[p] = bidx_part(g, starts[p], n_subs[p], n_bdys[p]);
    rowoff[p] = r;
    use_index = qidx[p] >= 0;
  }
  FAP_TRY(ensure_ws(ctx, ctx->ws_sub, (size_t)std::max<int64_t>(words, 1) * sizeof(int32_t)));
  std::vector<int32_t *> subs(count);
  for (int32_t p = 0; p < count; p++) subs[p] = (int32_t *)ctx->ws_sub.ptr + off[p];
  FwBatchDesc fd;
  build_fw_batch(count, subs.data(), n_subs, ld_sub.data(), &fd);
  // staging layout: sources | row pointers | closure jobs | closure segments
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_rp = up(4 * (size_t)nsrc);
  const size_t o_jobs = o_rp + up(sizeof(int32_t *) * (size_t)nsrc);
  const size_t o_segs = o_jobs + up(sizeof(fapk::BatchJob) * fd.jobs.size());
  const size_t total = o_segs + std::max<size_t>(sizeof(fapk::BatchSeg) * fd.segs.size(), 16);
  FAP_TRY(ensure_ws(ctx, ctx->ws_srcs, total));
  void *hs = nullptr;
  FAP_TRY(acquire_host_stage(ctx, total, &hs));
  char *h = (char *)hs;
  int32_t *h_src = (int32_t *)h;
  int32_t **h_rp = (int32_t **)(h + o_rp);
  for (int32_t p = 0, k = 0; p < count; p++)
    for (int32_t i = 0; i < n_bdys[p]; i++, k++) {
      h_src[k] = g->h_st2ed[starts[p] + i];
      h_rp[k] = rowblocks[p] + (int64_t)i * ld_rb;
    }
  memcpy(h + o_jobs, fd.jobs.data(), sizeof(fapk::BatchJob) * fd.jobs.size());
  if (!fd.segs.empty()) memcpy(h + o_segs, fd.segs.data(), sizeof(fapk::BatchSeg) * fd.segs.size());
  char *d = (char *)ctx->ws_srcs.ptr;
  FAP_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s));
  FAP_TRY(release_host_stage(ctx, s));
  if (sssp_after) FAP_HIP(hipStreamWaitEvent(s, sssp_after, 0));
  if (nsrc > 0 && use_index)
    FAP_TRY(bidx_rows(ctx, g, count, qidx.data(), rowoff.data(), (int32_t)nsrc,
                      (int32_t *const *)(d + o_rp), s));
  else if (nsrc > 0)
    FAP_TRY(run_sssp_rows(ctx, g, (const int32_t *)d, (int32_t)nsrc, nullptr, 0,
                          (int32_t *const *)(d + o_rp), s));
  if (sssp_done) FAP_HIP(hipEventRecord(sssp_done, s));
  for (int32_t p = 0; p < count; p++)
    FAP_TRY(launch_gather(g, starts[p], n_subs[p], n_bdys[p], rowblocks[p], ld_rb, subs[p],
                          ld_sub[p], s));
  FAP_TRY(launch_fw_batch(ctx, fd, (const fapk::BatchJob *)(d + o_jobs),
                          (const fapk::BatchSeg *)(d + o_segs), s));
  for (int32_t p = 0; p < count; p++) {
    const int32_t b = n_bdys[p];
    FAP_TRY(launch_minplus(subs[p] + (int64_t)b * ld_sub[p], ld_sub[p], rowblocks[p], ld_rb,
                           rowblocks[p] + (int64_t)b * ld_rb, ld_rb, n_subs[p] - b, g->V, b, 0,
                           s));
    FAP_TRY(launch_scatter(g, starts[p], n_subs[p], n_subs[p], subs[p], ld_sub[p], rowblocks[p],
                           ld_rb, s));
  }
  return FAP_OK;
}

// The boundary index is a per-graph cache: built (synchronously) by the
// first batched solve that may use it.
static int prepare_index(fap_ctx *ctx, const fap_graph_dev *g) {
  if (!ctx->opt_sssp_index || g->bidx_state != 0) return FAP_OK;
  return bidx_build(ctx, const_cast<fap_graph_dev *>(g));
}

int run_subgraph_solve_batch(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                             const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                             int32_t *const *rowblocks, int64_t ld_rb, hipStream_t s) {
  FAP_TRY(prepare_index(ctx, g));
  const int rc = enqueue_subgraph_batch(ctx, g, count, starts, n_subs, n_bdys, rowblocks, ld_rb, s);
  const hipError_t e = hipStreamSynchronize(s);
  if (rc != FAP_OK) return rc;
  if (e != hipSuccess) return hip_fail(e, "subgraph batch", __FILE__, __LINE__);
  return FAP_OK;
}

// Stage 1 alone for `count` parts: rows 0..b-1 of each part's row block
// (distances from its boundary vertices to every vertex), through the
// boundary index when it holds every part, else one SSSP launch over all
// sources.  Synchronous.
int run_boundary_rows(fap_ctx *ctx, const fap_graph_dev *g, int32_t count, const int32_t *starts,
                      const int32_t *n_subs, const int32_t *n_bdys, int32_t *const *rowblocks,
                      int64_t ld_rb, hipStream_t s) {
  FAP_TRY(prepare_index(ctx, g));
  int64_t nsrc = 0;
  for (int32_t p = 0; p < count; p++) nsrc += n_bdys[p];
  if (nsrc == 0) return FAP_OK;
  if (nsrc > INT32_MAX) return fail(FAP_E_ARG, "boundary rows: too many sources");
  std::vector<int32_t> qidx(count, -1), rowoff(count, 0);
  bool use_index = ctx->opt_sssp_index && g->bidx && (nsrc + 63) / 64 <= 65535;
  for (int32_t p = 0, r = 0; p < count && use_index; r += n_bdys[p], p++) {
    qidx[p] = bidx_part(g, starts[p], n_subs[p], n_bdys[p]);
    rowoff[p] = r;
    use_index = qidx[p] >= 0;
  }
  const size_t o_rp = (4 * (size_t)nsrc + 255) & ~(size_t)255;
  const size_t total = o_rp + sizeof(int32_t *) * (size_t)nsrc;
  FAP_TRY(ensure_ws(ctx, ctx->ws_srcs, total));
  void *hs = nullptr;
  FAP_TRY(acquire_host_stage(ctx, total, &hs));
  char *h = (char *)hs;
  int32_t *h_src = (int32_t *)h;
  int32_t **h_rp = (int32_t **)(h + o_rp);
  for (int32_t p = 0, k = 0; p < count; p++)
    for (int32_t i = 0; i < n_bdys[p]; i++, k++) {
      h_src[k] = g->h_st2ed[starts[p] + i];
      h_rp[k] = rowblocks[p] + (int64_t)i * ld_rb;
    }
  char *d = (char *)ctx->ws_srcs.ptr;
  FAP_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s));
  FAP_TRY(release_host_stage(ctx, s));
  if (use_index)
    FAP_TRY(bidx_rows(ctx, g, count, qidx.data(), rowoff.data(), (int32_t)nsrc,
                      (int32_t *const *)(d + o_rp), s));
  else
    FAP_TRY(run_sssp_rows(ctx, g, (const int32_t *)d, (int32_t)nsrc, nullptr, 0,
                          (int32_t *const *)(d + o_rp), s));
  FAP_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

// Groups of `group_parts` parts, alternating between ctx's stream and its
// peer context's: one group's stage 1 (boundary index iterations, or SSSP)
// runs beside the other's closure and products (VALU-bound).  Each group's
// stage 1 waits for the previous group's: two SSSP launches side by side
// took longer than one after the other (2,048 concurrent sources thrash the
// caches).  Synchronous.
int run_subgraph_groups(fap_ctx *ctx, const fap_graph_dev *g, int32_t count, const int32_t *starts,
                        const int32_t *n_subs, const int32_t *n_bdys, int32_t *const *rowblocks,
                        int64_t ld_rb, int32_t group_parts) {
  if (count <= 0) return FAP_OK;
  if (count <= group_parts)
    return run_subgraph_solve_batch(ctx, g, count, starts, n_subs, n_bdys, rowblocks, ld_rb,
                                    ctx->stream);
  if (!ctx->peer) {
    FAP_TRY(fap_ctx_create(ctx->device, &ctx->peer));
  }
  FAP_TRY(prepare_index(ctx, g));
  fap_ctx *peer = ctx->peer;
  peer->opt_sssp_index = ctx->opt_sssp_index;
  peer->opt_bidx_delta = ctx->opt_bidx_delta;
  peer->opt_bidx_tile = ctx->opt_bidx_tile;
  peer->opt_fw_lookahead = ctx->opt_fw_lookahead;
  peer->opt_fw_tile = ctx->opt_fw_tile;
  peer->opt_sssp_threads = ctx->opt_sssp_threads;
  peer->opt_sssp_delta = ctx->opt_sssp_delta;
  peer->opt_sssp_group = ctx->opt_sssp_group;
  peer->opt_sssp_bin = ctx->opt_sssp_bin;
  peer->opt_sssp_pack = ctx->opt_sssp_pack;
  peer->opt_sssp_wgs = ctx->opt_sssp_wgs;
  peer->opt_sssp_arcs = ctx->opt_sssp_arcs;
  peer->opt_batch_grid = ctx->opt_batch_grid;
  peer->tune = ctx->tune;
  fap_ctx *c[2] = {ctx, peer};
  int rc = FAP_OK;
  for (int32_t first = 0, gi = 0; first < count && rc == FAP_OK; first += group_parts, gi++) {
    fap_ctx *cc = c[gi & 1];
    const int32_t n = std::min(group_parts, count - first);
    rc = enqueue_subgraph_batch(cc, g, n, starts + first, n_subs + first, n_bdys + first,
                                rowblocks + first, ld_rb, cc->stream,
                                gi > 0 ? c[(gi - 1) & 1]->ev_sssp : nullptr, cc->ev_sssp);
  }
  const hipError_t e1 = hipStreamSynchronize(ctx->stream);
  const hipError_t e2 = hipStreamSynchronize(peer->stream);
  if (rc != FAP_OK) return rc;
  if (e1 != hipSuccess) return hip_fail(e1, "subgraph groups", __FILE__, __LINE__);
  if (e2 != hipSuccess) return hip_fail(e2, "subgraph groups", __FILE__, __LINE__);
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_subgraph_solve_batch(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                             const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                             int32_t *const *rowblocks, int64_t ld_rb, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (count < 0 || (count > 0 && (!starts || !n_subs || !n_bdys || !rowblocks)))
    return fail(FAP_E_ARG, "subgraph batch: bad arguments");
  if (ld_rb < g->V) return fail(FAP_E_ARG, "subgraph batch: ld_rb < V");
  for (int32_t p = 0; p < count; p++) {
    FAP_TRY(check_part(g, starts[p], n_subs[p], n_bdys[p]));
    if (n_subs[p] > 0 && !rowblocks[p]) return fail(FAP_E_ARG, "subgraph batch: null row block");
  }
  return run_subgraph_solve_batch(ctx, g, count, starts, n_subs, n_bdys, rowblocks, ld_rb,
                                  pick_stream(ctx, stream));
}

int fap_subgraph_boundary_rows(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                               const int32_t *starts, const int32_t *n_subs,
                               const int32_t *n_bdys, int32_t *const *rowblocks, int64_t ld_rb,
                               void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (count < 0 || (count > 0 && (!starts || !n_subs || !n_bdys || !rowblocks)))
    return fail(FAP_E_ARG, "boundary rows: bad arguments");
  if (ld_rb < g->V) return fail(FAP_E_ARG, "boundary rows: ld_rb < V");
  if ((int64_t)g->h_st2ed.size() != g->V)
    return fail(FAP_E_STATE, "graph has no partition (fap_graph_set_partition)");
  for (int32_t p = 0; p < count; p++) {
    FAP_TRY(check_part(g, starts[p], n_subs[p], n_bdys[p]));
    if (n_bdys[p] > 0 && !rowblocks[p]) return fail(FAP_E_ARG, "boundary rows: null row block");
  }
  return run_boundary_rows(ctx, g, count, starts, n_subs, n_bdys, rowblocks, ld_rb,
                           pick_stream(ctx, stream));
}

int fap_subgraph_solve_groups(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                              const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                              int32_t *const *rowblocks, int64_t ld_rb, int32_t group_parts) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (count < 0 || (count > 0 && (!starts || !n_subs || !n_bdys || !rowblocks)))
    return fail(FAP_E_ARG, "subgraph groups: bad arguments");
  if (group_parts < 1) return fail(FAP_E_ARG, "subgraph groups: group_parts < 1");
  if (ld_rb < g->V) return fail(FAP_E_ARG, "subgraph groups: ld_rb < V");
  for (int32_t p = 0; p < count; p++) {
    FAP_TRY(check_part(g, starts[p], n_subs[p], n_bdys[p]));
    if (n_subs[p] > 0 && !rowblocks[p]) return fail(FAP_E_ARG, "subgraph groups: null row block");
  }
  return run_subgraph_groups(ctx, g, count, starts, n_subs, n_bdys, rowblocks, ld_rb, group_parts);
}

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
