// dense.hip -- dense distance matrices of whole graphs (K = 1 jobs).
//
// For a standalone graph the reference pipeline with K = 1 has no boundary
// vertices: stage 1 is empty and stage 2 closes the matrix built by
// MysubMatBuild_path (fap/utils/util_Myalgorithm.h:30-73) from all arcs --
// INF, then every arc in CSR order (the last duplicate wins, :50), then a
// zero diagonal (:67-72).  These kernels build that matrix for one graph or
// for a batch of graphs in one pair of launches.
#include <algorithm>
#include <vector>

#include "batch_desc.h"
#include "common.h"

namespace fapk {

__device__ __forceinline__ int find_job(const DenseJob *jobs, int count, int64_t x, bool rows) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b = rows ? (int64_t)jobs[mid].row_base : jobs[mid].cell_base;
    if (b <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void dense_fill_kernel(const DenseJob *__restrict__ jobs, int count, int64_t cells) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < cells; x += stride) {
    const DenseJob &jb = jobs[find_job(jobs, count, x, false)];
    const int64_t local = x - jb.cell_base;
    const int i = (int)(local / jb.n), j = (int)(local - (local / jb.n) * jb.n);
    jb.mat[(int64_t)i * jb.ld + j] = (i == j) ? 0 : 0x3FFFFFFF;
  }
}

__global__ void dense_arcs_kernel(const DenseJob *__restrict__ jobs, int count, int32_t rows) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= rows) return;
  const DenseJob &jb = jobs[find_job(jobs, count, x, true)];
  const int u = x - jb.row_base;
  int32_t *row = jb.mat + (int64_t)u * jb.ld;
  for (int e = jb.row_offset[u]; e < jb.row_offset[u + 1]; e++) {
    const int2 ed = jb.edges[e];
    row[ed.x] = ed.y;
  }
  row[u] = 0;
}

}  // namespace fapk

namespace fapi {

int launch_dense_jobs(fap_ctx *ctx, const fapk::DenseJob *d_jobs, int32_t njobs, int64_t cells,
                      int64_t rows, hipStream_t s) {
  if (njobs <= 0 || cells <= 0) return FAP_OK;
  if (rows > INT32_MAX) return fail(FAP_E_ARG, "dense build: too many rows");
  const int64_t blocks = std::min<int64_t>((cells + 255) / 256, (int64_t)ctx->num_cus * 32);
  hipLaunchKernelGGL(fapk::dense_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_jobs,
                     njobs, cells);
  FAP_CHECK_LAUNCH();
  hipLaunchKernelGGL(fapk::dense_arcs_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0,
                     s, d_jobs, njobs, (int32_t)rows);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int run_dense_from_graphs(fap_ctx *ctx, int32_t count, const fap_graph_dev *const *graphs,
                          int32_t *const *mats, const int64_t *ld, hipStream_t s) {
  std::vector<fapk::DenseJob> jobs;
  jobs.reserve(count);
  int64_t cells = 0;
  int64_t rows = 0;
  for (int32_t j = 0; j < count; j++) {
    const fap_graph_dev *g = graphs[j];
    if (g->V == 0) continue;
    jobs.push_back(fapk::DenseJob{g->row_offset, g->edges, mats[j], ld[j], cells, (int32_t)rows,
                                  g->V});
    cells += (int64_t)g->V * g->V;
    rows += g->V;
  }
  if (jobs.empty()) return FAP_OK;
  if (rows > INT32_MAX) return fail(FAP_E_ARG, "dense build: too many rows");
  const size_t bytes = sizeof(fapk::DenseJob) * jobs.size();
  FAP_TRY(ensure_ws(ctx, ctx->ws_dense, bytes));
  FAP_HIP(hipMemcpyAsync(ctx->ws_dense.ptr, jobs.data(), bytes, hipMemcpyHostToDevice, s));
  FAP_TRY(launch_dense_jobs(ctx, (const fapk::DenseJob *)ctx->ws_dense.ptr, (int32_t)jobs.size(),
                            cells, rows, s));
  // the descriptor copy reads `jobs` asynchronously
  FAP_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_dense_from_graphs(fap_ctx *ctx, int32_t count,
                                     const fap_graph_dev *const *graphs, int32_t *const *d_mats,
                                     const int64_t *ld, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (count < 0 || (count > 0 && (!graphs || !d_mats || !ld)))
    return fail(FAP_E_ARG, "dense build: bad arguments");
  for (int32_t j = 0; j < count; j++) {
    if (!graphs[j]) return fail(FAP_E_ARG, "dense build: null graph");
    if (graphs[j]->V > 0 && (!d_mats[j] || ld[j] < graphs[j]->V))
      return fail(FAP_E_ARG, "dense build: bad matrix/stride for job " + std::to_string(j));
  }
  return run_dense_from_graphs(ctx, count, graphs, d_mats, ld, pick_stream(ctx, stream));
}
