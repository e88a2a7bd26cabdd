// ops_f32.cpp -- the reference's host-memory float operators on the GPU
// (include/fap_ops.h).  Each call: convert float -> int32 (>= 1e8 is INF),
// copy to HBM, run the int32 kernel, copy back, convert int32 -> float
// (INF -> 1e8f).  The device buffers live for the call only, matching the
// reference's per-call allocation (fap/kernel/hip/minplus.cpp:91-128).
// Path matrices, when the caller passes them, are written with the
// reference's meaning: Dijkstra parents (fap/kernel/batch_sssp.h:82-83,110),
// FW hop labels path[i][j] = path[k][j] (fap/kernel/floyd.h:44) and the
// min-plus label mat2_path[k][j] of the first minimising k
// (fap/kernel/minplus.h:57-60).
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "fap_ops.h"

namespace fapi {

int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);
int launch_minplus(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb, int32_t *C,
                   int64_t ldc, int32_t M, int32_t N, int32_t K, int accumulate, hipStream_t s);
int run_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
             int32_t *d_out, int64_t ld_out, hipStream_t s);
int run_pred_rows(fap_ctx *ctx, fap_graph_dev *g, const int32_t *d_dist, int64_t ld,
                  const int32_t *d_own, int32_t nrows, int32_t *d_pred, int64_t ldp,
                  hipStream_t s);
int run_minplus_path(const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb,
                     const int32_t *Bp, int64_t ldbp, int32_t *C, int64_t ldc, int32_t *Cp,
                     int64_t ldcp, int32_t M, int32_t N, int32_t K, hipStream_t s);
int run_fw_path(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, int32_t *p, int64_t ldp,
                hipStream_t s);

namespace {

constexpr float kRefInf = 1e8f;
constexpr float kExact = 16777216.f;  // 2^24

std::mutex g_mu;
fap_ctx *g_ctx = nullptr;

int default_ctx(fap_ctx **out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ctx) FAP_TRY(fap_ctx_create(0, &g_ctx));
  *out = g_ctx;
  return FAP_OK;
}

int to_i32(const float *src, int64_t n, std::vector<int32_t> *dst, const char *what) {
  dst->resize((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    const float x = src[i];
    if (x >= kRefInf) {
      (*dst)[i] = kInf;
    } else if (!(x >= 0.f) || x != floorf(x) || x >= kExact) {
      return fail(FAP_E_ARG, std::string(what) + ": value " + std::to_string(x) + " at " +
                                 std::to_string(i) +
                                 " is not a non-negative integer below 2^24 (or >= 1e8 for INF)");
    } else {
      (*dst)[i] = (int32_t)x;
    }
  }
  return FAP_OK;
}

void to_f32(const std::vector<int32_t> &src, float *dst) {
  for (size_t i = 0; i < src.size(); i++) dst[i] = src[i] >= kInf ? kRefInf : (float)src[i];
}

// RAII device buffer on the default context
struct DevBuf {
  fap_ctx *ctx;
  void *p = nullptr;
  explicit DevBuf(fap_ctx *c) : ctx(c) {}
  int alloc(size_t bytes) { return fap_dev_alloc(ctx, bytes, &p); }
  ~DevBuf() {
    if (p) fap_dev_free(ctx, p);
  }
};

}  // namespace
}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_handle_boundry_path(float *subGraph, int *subGraph_path, int vertexs, int edges,
                            int bdy_num, int *adj_size, int *row_offset, int *col_val,
                            float *weights, int *st2ed, int offset) {
  (void)adj_size;
  if (bdy_num == 0) return FAP_OK;
  if (!subGraph || !row_offset || !st2ed || vertexs <= 0 || edges < 0 || bdy_num < 0 ||
      offset < 0 || (int64_t)offset + bdy_num > vertexs || (edges > 0 && (!col_val || !weights)))
    return fail(FAP_E_ARG, "handle_boundry_path: bad arguments");
  if (row_offset[vertexs] != edges)
    return fail(FAP_E_ARG, "handle_boundry_path: row_offset[vertexs] != edges");
  std::vector<int32_t> w;
  FAP_TRY(to_i32(weights, edges, &w, "handle_boundry_path weight"));
  fap_ctx *ctx = nullptr;
  FAP_TRY(default_ctx(&ctx));
  fap_graph_dev *g = nullptr;
  FAP_TRY(fap_graph_upload(ctx, vertexs, edges, row_offset, col_val, w.data(), &g));
  int rc = FAP_OK;
  {
    DevBuf src(ctx), out(ctx), pred(ctx);
    rc = src.alloc(sizeof(int32_t) * bdy_num);
    if (rc == FAP_OK) rc = out.alloc(sizeof(int32_t) * (size_t)bdy_num * vertexs);
    if (rc == FAP_OK && subGraph_path)
      rc = pred.alloc(sizeof(int32_t) * (size_t)bdy_num * vertexs);
    if (rc == FAP_OK) {
      for (int i = 0; i < bdy_num && rc == FAP_OK; i++)
        if (st2ed[offset + i] < 0 || st2ed[offset + i] >= vertexs)
          rc = fail(FAP_E_ARG, "handle_boundry_path: st2ed entry out of range");
    }
    if (rc == FAP_OK) rc = fap_memcpy_h2d(ctx, src.p, st2ed + offset, sizeof(int32_t) * bdy_num);
    if (rc == FAP_OK)
      rc = run_sssp(ctx, g, (const int32_t *)src.p, bdy_num, (int32_t *)out.p, vertexs,
                    ctx->stream);
    if (rc == FAP_OK && subGraph_path)
      rc = run_pred_rows(ctx, g, (const int32_t *)out.p, vertexs, (const int32_t *)src.p, bdy_num,
                         (int32_t *)pred.p, vertexs, ctx->stream);
    if (rc == FAP_OK) {
      std::vector<int32_t> h((size_t)bdy_num * vertexs);
      rc = fap_memcpy_d2h(ctx, h.data(), out.p, h.size() * sizeof(int32_t));
      if (rc == FAP_OK) to_f32(h, subGraph);
    }
    if (rc == FAP_OK && subGraph_path)
      rc = fap_memcpy_d2h(ctx, subGraph_path, pred.p,
                          sizeof(int32_t) * (size_t)bdy_num * vertexs);
  }
  fap_graph_free(ctx, g);
  return rc;
}

int fap_floyd_path(int num, float *mat, int *path) {
  if (num == 0) return FAP_OK;
  if (num < 0 || !mat) return fail(FAP_E_ARG, "floyd_path: bad arguments");
  const int64_t n2 = (int64_t)num * num;
  std::vector<int32_t> m;
  FAP_TRY(to_i32(mat, n2, &m, "floyd_path"));
  fap_ctx *ctx = nullptr;
  FAP_TRY(default_ctx(&ctx));
  DevBuf d(ctx);
  FAP_TRY(d.alloc(sizeof(int32_t) * n2));
  FAP_TRY(fap_memcpy_h2d(ctx, d.p, m.data(), sizeof(int32_t) * n2));
  if (path) {
    DevBuf dp(ctx);
    FAP_TRY(dp.alloc(sizeof(int32_t) * n2));
    FAP_TRY(fap_memcpy_h2d(ctx, dp.p, path, sizeof(int32_t) * n2));
    FAP_TRY(run_fw_path(ctx, (int32_t *)d.p, num, num, (int32_t *)dp.p, num, ctx->stream));
    FAP_TRY(fap_memcpy_d2h(ctx, path, dp.p, sizeof(int32_t) * n2));
  } else {
    FAP_TRY(launch_fw(ctx, (int32_t *)d.p, num, num, ctx->stream));
  }
  FAP_TRY(fap_memcpy_d2h(ctx, m.data(), d.p, sizeof(int32_t) * n2));
  to_f32(m, mat);
  return FAP_OK;
}

int fap_min_plus_path_advanced(float *mat1, float *mat2, int *mat2_path, float *res, int *path,
                               int M, int N, int K) {
  if (M < 0 || N < 0 || K < 0) return fail(FAP_E_ARG, "min_plus_path_advanced: negative size");
  if (M == 0 || N == 0) return FAP_OK;
  if (!res || (K > 0 && (!mat1 || !mat2)))
    return fail(FAP_E_ARG, "min_plus_path_advanced: null operand");
  if (path && K > 0 && !mat2_path)
    return fail(FAP_E_ARG, "min_plus_path_advanced: path requested without mat2_path");
  std::vector<int32_t> a, b, c((size_t)M * N);
  FAP_TRY(to_i32(mat1, (int64_t)M * K, &a, "min_plus mat1"));
  FAP_TRY(to_i32(mat2, (int64_t)K * N, &b, "min_plus mat2"));
  fap_ctx *ctx = nullptr;
  FAP_TRY(default_ctx(&ctx));
  DevBuf da(ctx), db(ctx), dc(ctx);
  FAP_TRY(da.alloc(sizeof(int32_t) * a.size()));
  FAP_TRY(db.alloc(sizeof(int32_t) * b.size()));
  FAP_TRY(dc.alloc(sizeof(int32_t) * c.size()));
  FAP_TRY(fap_memcpy_h2d(ctx, da.p, a.data(), sizeof(int32_t) * a.size()));
  FAP_TRY(fap_memcpy_h2d(ctx, db.p, b.data(), sizeof(int32_t) * b.size()));
  if (path) {
    // the first minimising k fixes the label: one kernel computes both
    DevBuf dbp(ctx), dcp(ctx);
    if (K > 0) {
      FAP_TRY(dbp.alloc(sizeof(int32_t) * b.size()));
      FAP_TRY(fap_memcpy_h2d(ctx, dbp.p, mat2_path, sizeof(int32_t) * b.size()));
    }
    FAP_TRY(dcp.alloc(sizeof(int32_t) * c.size()));
    FAP_TRY(run_minplus_path((const int32_t *)da.p, K, (const int32_t *)db.p, N,
                             (const int32_t *)dbp.p, N, (int32_t *)dc.p, N, (int32_t *)dcp.p, N,
                             M, N, K, ctx->stream));
    FAP_TRY(fap_memcpy_d2h(ctx, path, dcp.p, sizeof(int32_t) * c.size()));
  } else {
    FAP_TRY(launch_minplus((const int32_t *)da.p, K, (const int32_t *)db.p, N, (int32_t *)dc.p,
                           N, M, N, K, 0, ctx->stream));
  }
  FAP_TRY(fap_memcpy_d2h(ctx, c.data(), dc.p, sizeof(int32_t) * c.size()));
  to_f32(c, res);
  return FAP_OK;
}

}  // extern "C"
