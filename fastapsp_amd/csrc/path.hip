// path.hip -- predecessor rows from exact distance rows (is_path_needed).
//
// The reference carries a path matrix beside every distance matrix: the
// Dijkstra parent (fap/kernel/batch_sssp.h:82-83,110), the FW update
// path[i][j] = path[k][j] (fap/kernel/floyd.h:44), the min-plus update
// path = B_path[k][j] (fap/kernel/minplus.h:60), the part matrix's arcs and
// diagonal (fap/utils/util_Myalgorithm.h:50-51,71).  In every case
// path[i][j] is the vertex before j on a shortest path from row i's vertex
// to j, the row's own vertex for j itself, and -1 when j is unreachable.
//
// Here the distances are computed first (exact, unique) and the predecessor
// is read off them afterwards: u is a valid predecessor of j on row i iff an
// arc u -> j of weight w has d[i][u] + w == d[i][j].  One pass picks, per
// cell, the smallest such u with w > 0 (so d[i][u] < d[i][j]; following
// predecessors strictly decreases the distance and ends at the row's vertex).
// Cells whose only tight arcs have weight 0 (a plateau of equal distance) are
// marked and resolved in further passes: a marked cell takes the smallest
// tight zero-weight in-neighbour resolved in an earlier pass, so walks
// through a plateau end at a vertex resolved earlier, and the rows are a
// function of the distances alone (the same on every run).  Which valid predecessor is
// chosen when several exist is not fixed by the reference (its CPU, CUDA and
// HIP paths differ); validity is.
//
// In-arcs come from the transposed CSR, built once per graph on first use.
#include <algorithm>
#include <vector>

#include "common.h"

namespace fapk {

constexpr int32_t kUnresolved = -2;

__global__ void pred_kernel(const int32_t *__restrict__ dist, int64_t ld,
                            const int32_t *__restrict__ own, int32_t nrows, int32_t V,
                            const int32_t *__restrict__ in_off, const int2 *__restrict__ in_arcs,
                            int32_t *__restrict__ pred, int64_t ldp,
                            unsigned int *__restrict__ unresolved) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= V) return;
  const int e0 = in_off[j], e1 = in_off[j + 1];
  unsigned int marked = 0;
  for (int i = blockIdx.y; i < nrows; i += gridDim.y) {
    const int32_t *d = dist + (int64_t)i * ld;
    const int dj = d[j];
    const int me = own[i];
    int p;
    if (j == me) {
      p = me;
    } else if (dj >= 0x3FFFFFFF) {
      p = -1;
    } else {
      p = 0x7FFFFFFF;
      bool plateau = false;
      for (int e = e0; e < e1; e++) {
        const int2 a = in_arcs[e];  // {source vertex, weight}
        const int du = d[a.x];
        if (du + a.y == dj) {
          if (a.y > 0) p = min(p, a.x);
          else plateau = true;
        }
      }
      if (p == 0x7FFFFFFF) {
        // no tight arc with a positive weight: a plateau cell, resolved by
        // the passes below (with no tight arc at all, `plateau` is false and
        // the passes report the rows as inconsistent)
        p = kUnresolved;
        marked++;
        (void)plateau;
      }
    }
    pred[(int64_t)i * ldp + j] = p;
  }
  if (marked) atomicAdd(unresolved, marked);
}

// One resolution pass over plateau cells: take the smallest tight
// zero-weight in-neighbour resolved BEFORE this pass.  A cell resolved here
// is written as -(u + 3) (readers of the same pass skip it) and made
// positive by pred_settle_kernel after the pass, so the choice does not
// depend on the order in which workgroups run.  Counts the cells resolved.
__global__ void pred_plateau_kernel(const int32_t *__restrict__ dist, int64_t ld, int32_t nrows,
                                    int32_t V, const int32_t *__restrict__ in_off,
                                    const int2 *__restrict__ in_arcs, int32_t *pred, int64_t ldp,
                                    unsigned int *__restrict__ resolved) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= V) return;
  const int e0 = in_off[j], e1 = in_off[j + 1];
  unsigned int n = 0;
  for (int i = blockIdx.y; i < nrows; i += gridDim.y) {
    int32_t *pr = pred + (int64_t)i * ldp;
    if (pr[j] != kUnresolved) continue;
    const int32_t *d = dist + (int64_t)i * ld;
    const int dj = d[j];
    int p = 0x7FFFFFFF;
    for (int e = e0; e < e1; e++) {
      const int2 a = in_arcs[e];
      if (a.y == 0 && d[a.x] == dj &&
          __hip_atomic_load(pr + a.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= -1)
        p = min(p, a.x);
    }
    if (p != 0x7FFFFFFF) {
      __hip_atomic_store(pr + j, -(p + 3), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      n++;
    }
  }
  if (n) atomicAdd(resolved, n);
}

// After a plateau pass: cells resolved in it (-(u + 3)) become u.
__global__ void pred_settle_kernel(int32_t *pred, int64_t ldp, int32_t nrows, int32_t V) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= V) return;
  for (int i = blockIdx.y; i < nrows; i += gridDim.y) {
    int32_t *c = pred + (int64_t)i * ldp + j;
    const int32_t v = *c;
    if (v <= -3) *c = -v - 3;
  }
}

}  // namespace fapk

namespace fapi {

// Transposed CSR of g (in-arcs of every vertex as {source, weight}, sources
// ascending, arcs of one source in CSR order), built on the host from the
// device copy.
static int ensure_in_csr(fap_ctx *ctx, fap_graph_dev *g, hipStream_t s) {
  if (g->in_offset) return FAP_OK;
  const int32_t V = g->V;
  const int64_t E = g->E;
  std::vector<int32_t> ro(V + 1);
  std::vector<int2> ed((size_t)E);
  FAP_HIP(hipMemcpyAsync(ro.data(), g->row_offset, sizeof(int32_t) * (V + 1),
                         hipMemcpyDeviceToHost, s));
  if (E) FAP_HIP(hipMemcpyAsync(ed.data(), g->edges, sizeof(int2) * E, hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> off(V + 1, 0);
  for (int64_t e = 0; e < E; e++) off[ed[e].x + 1]++;
  for (int32_t v = 0; v < V; v++) off[v + 1] += off[v];
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  std::vector<int2> in((size_t)E);
  for (int32_t u = 0; u < V; u++)
    for (int32_t e = ro[u]; e < ro[u + 1]; e++) in[fill[ed[e].x]++] = make_int2(u, ed[e].y);
  int32_t *d_off = nullptr;
  int2 *d_in = nullptr;
  hipError_t e1 = hipMalloc(&d_off, sizeof(int32_t) * (V + 1));
  hipError_t e2 = hipMalloc(&d_in, sizeof(int2) * std::max<int64_t>(E, 1));
  if (e1 != hipSuccess || e2 != hipSuccess) {
    if (d_off) (void)hipFree(d_off);
    if (d_in) (void)hipFree(d_in);
    (void)hipGetLastError();
    return fail(FAP_E_NOMEM, "transposed CSR allocation failed");
  }
  hipError_t e = hipMemcpyAsync(d_off, off.data(), sizeof(int32_t) * (V + 1),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess && E)
    e = hipMemcpyAsync(d_in, in.data(), sizeof(int2) * E, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d_off);
    (void)hipFree(d_in);
    return hip_fail(e, "transposed CSR copy", __FILE__, __LINE__);
  }
  g->in_offset = d_off;
  g->in_arcs = d_in;
  (void)ctx;
  return FAP_OK;
}

int run_pred_rows(fap_ctx *ctx, fap_graph_dev *g, const int32_t *d_dist, int64_t ld,
                  const int32_t *d_own, int32_t nrows, int32_t *d_pred, int64_t ldp,
                  hipStream_t s) {
  if (nrows <= 0 || g->V <= 0) return FAP_OK;
  FAP_TRY(ensure_in_csr(ctx, g, s));
  FAP_TRY(ensure_ws(ctx, ctx->ws_misc, 64));
  unsigned int *cnt = (unsigned int *)ctx->ws_misc.ptr;
  const int V = g->V;
  dim3 blk(256), grd((V + 255) / 256, std::min(nrows, 256));
  FAP_TRY(launch_fill_u32(cnt, 1, 0u, s));
  hipLaunchKernelGGL(fapk::pred_kernel, grd, blk, 0, s, d_dist, ld, d_own, nrows, V,
                     g->in_offset, g->in_arcs, d_pred, ldp, cnt);
  FAP_CHECK_LAUNCH();
  unsigned int h = 0;
  FAP_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  int64_t left = h;
  // each pass resolves at least one plateau cell per row that has any
  // (some unresolved cell of a plateau always has a resolved tight
  // in-neighbour), so V passes bound the loop
  for (int pass = 0; left > 0; pass++) {
    if (pass >= V) return fail(FAP_E_STATE, "predecessor plateaus did not resolve");
    FAP_TRY(launch_fill_u32(cnt, 1, 0u, s));
    hipLaunchKernelGGL(fapk::pred_plateau_kernel, grd, blk, 0, s, d_dist, ld, nrows, V,
                       g->in_offset, g->in_arcs, d_pred, ldp, cnt);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(fapk::pred_settle_kernel, grd, blk, 0, s, d_pred, ldp, nrows, V);
    FAP_CHECK_LAUNCH();
    FAP_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    FAP_HIP(hipStreamSynchronize(s));
    if (h == 0) return fail(FAP_E_STATE, "predecessor plateaus made no progress");
    left -= h;
  }
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_pred_rows(fap_ctx *ctx, fap_graph_dev *g, const int32_t *d_dist, int64_t ld,
                             const int32_t *d_row_vertex, int32_t nrows, int32_t *d_pred,
                             int64_t ld_pred, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (nrows < 0) return fail(FAP_E_ARG, "pred: negative row count");
  if (nrows == 0) return FAP_OK;
  if (!d_dist || !d_row_vertex || !d_pred || ld < g->V || ld_pred < g->V)
    return fail(FAP_E_ARG, "pred: bad pointers or strides");
  return run_pred_rows(ctx, g, d_dist, ld, d_row_vertex, nrows, d_pred, ld_pred,
                       pick_stream(ctx, stream));
}

extern "C" int fap_subgraph_pred(fap_ctx *ctx, fap_graph_dev *g, int32_t start, int32_t n_sub,
                                 const int32_t *rowblock, int64_t ld_rb, int32_t *predblock,
                                 int64_t ld_pred, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (!g->st2ed) return fail(FAP_E_STATE, "graph has no partition (fap_graph_set_partition)");
  if (start < 0 || n_sub < 0 || (int64_t)start + n_sub > g->V)
    return fail(FAP_E_ARG, "part range outside the vertex permutation");
  if (n_sub == 0) return FAP_OK;
  if (!rowblock || !predblock || ld_rb < g->V || ld_pred < g->V)
    return fail(FAP_E_ARG, "pred: bad pointers or strides");
  return run_pred_rows(ctx, g, rowblock, ld_rb, g->st2ed + start, n_sub, predblock, ld_pred,
                       pick_stream(ctx, stream));
}
