// ctx.hip -- context, errors, device memory helpers, resident graph.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "common.h"

namespace fapi {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e,
           hipGetErrorString(e), what, file, line);
  g_last_error = buf;
  return FAP_E_HIP;
}

int ensure_ws(fap_ctx *ctx, Workspace &ws, size_t bytes) {
  if (ws.bytes >= bytes) return FAP_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ws.bytes >= bytes) return FAP_OK;
  if (ws.ptr) {
    FAP_HIP(hipDeviceSynchronize());
    FAP_HIP(hipFree(ws.ptr));
    ws.ptr = nullptr;
    ws.bytes = 0;
  }
  size_t want = std::max(bytes, (size_t)256);
  want = (want + 4095) & ~(size_t)4095;
  hipError_t e = hipMalloc(&ws.ptr, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ws.ptr = nullptr;
    return fail(FAP_E_NOMEM, "workspace allocation of " + std::to_string(want) +
                                 " bytes failed: " + hipGetErrorString(e));
  }
  ws.bytes = want;
  return FAP_OK;
}

int acquire_host_stage(fap_ctx *ctx, size_t bytes, void **out) {
  if (ctx->stage_pending) {
    FAP_HIP(hipEventSynchronize(ctx->ev_stage));
    ctx->stage_pending = false;
  }
  if (ctx->host_stage_bytes < bytes) {
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    ctx->host_stage = nullptr;
    ctx->host_stage_bytes = 0;
    const size_t want = std::max<size_t>((bytes + 4095) & ~(size_t)4095, 1 << 16);
    hipError_t e = hipHostMalloc(&ctx->host_stage, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ctx->host_stage = nullptr;
      return fail(FAP_E_NOMEM, "pinned staging allocation failed");
    }
    ctx->host_stage_bytes = want;
  }
  *out = ctx->host_stage;
  return FAP_OK;
}

int release_host_stage(fap_ctx *ctx, hipStream_t s) {
  FAP_HIP(hipEventRecord(ctx->ev_stage, s));
  ctx->stage_pending = true;
  return FAP_OK;
}

// Grid-stride fill; used instead of hipMemsetAsync so that the sequence can
// be captured in a graph.
__global__ void fill_i32_kernel(int32_t *__restrict__ d, int64_t n, int32_t v) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) d[i] = v;
}

int launch_fill_i32(int32_t *d, int64_t n, int32_t value, hipStream_t s) {
  if (n <= 0) return FAP_OK;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n, value);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// Plain streaming copy (16-byte vectors when both pointers allow it); also
// the known-size reference kernel for calibrating HBM byte counters.
__global__ void copy_i32_kernel(const int32_t *__restrict__ src, int32_t *__restrict__ dst,
                                int64_t n, int vec) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vec) {
    const int64_t n4 = n >> 2;
    const int4 *s4 = reinterpret_cast<const int4 *>(src);
    int4 *d4 = reinterpret_cast<int4 *>(dst);
    for (int64_t i = t; i < n4; i += stride) d4[i] = s4[i];
    for (int64_t i = (n4 << 2) + t; i < n; i += stride) dst[i] = src[i];
  } else {
    for (int64_t i = t; i < n; i += stride) dst[i] = src[i];
  }
}

int launch_fill_u32(uint32_t *d, int64_t n, uint32_t value, hipStream_t s) {
  return launch_fill_i32((int32_t *)d, n, (int32_t)value, s);
}

// Wrapping sums: plain and position-weighted ((i+1) * value), per block then
// one atomic add per block.
__global__ void checksum_kernel(const uint32_t *__restrict__ d, int64_t n,
                                unsigned long long *out) {
  unsigned long long s = 0, ws = 0;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    unsigned long long v = d[i];
    s += v;
    ws += v * (unsigned long long)(i + 1);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    ws += __shfl_down(ws, off);
  }
  __shared__ unsigned long long ps[16], pw[16];
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    ps[wv] = s;
    pw[wv] = ws;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      a += ps[i];
      b += pw[i];
    }
    atomicAdd(out, a);
    atomicAdd(out + 1, b);
  }
}

}  // namespace fapi

using namespace fapi;

extern "C" {

const char *fap_last_error(void) { return g_last_error.c_str(); }
const char *fap_version(void) { return "fap-mi355x 0.1 (gfx950)"; }

int fap_device_count(int *count) {
  if (!count) return fail(FAP_E_ARG, "count is null");
  FAP_HIP(hipGetDeviceCount(count));
  return FAP_OK;
}

int fap_ctx_create(int device, fap_ctx **out) {
  if (!out) return fail(FAP_E_ARG, "out is null");
  int n = 0;
  FAP_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(FAP_E_ARG, "device " + std::to_string(device) + " out of range (" +
                               std::to_string(n) + " devices)");
  FAP_HIP(hipSetDevice(device));
  fap_ctx *c = new fap_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipGetDeviceProperties", __FILE__, __LINE__);
  }
  c->num_cus = prop.multiProcessorCount;
  c->lds_per_block = prop.sharedMemPerBlock;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->host_pinned, 4096, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_poll, hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fw_la, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fw_panel, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_stage, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_sssp, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "context setup", __FILE__, __LINE__);
  }
  c->tune = tune_env_cache();
  *out = c;
  return FAP_OK;
}

int fap_ctx_destroy(fap_ctx *ctx) {
  if (!ctx) return FAP_OK;
  if (ctx->peer) (void)fap_ctx_destroy(ctx->peer);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  Workspace *all[] = {&ctx->ws_sub, &ctx->ws_queue, &ctx->ws_stamp, &ctx->ws_small,
                      &ctx->ws_misc, &ctx->ws_batch, &ctx->ws_dense, &ctx->ws_srcs,
                      &ctx->ws_bidx, &ctx->ws_bbf};
  for (Workspace *w : all)
    if (w->ptr) (void)hipFree(w->ptr);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
  if (ctx->host_bbf) (void)hipHostFree(ctx->host_bbf);
  if (ctx->ev_stage) (void)hipEventDestroy(ctx->ev_stage);
  if (ctx->ev_sssp) (void)hipEventDestroy(ctx->ev_sssp);
  if (ctx->ev_poll) (void)hipEventDestroy(ctx->ev_poll);
  if (ctx->ev_fw_la) (void)hipEventDestroy(ctx->ev_fw_la);
  if (ctx->ev_fw_panel) (void)hipEventDestroy(ctx->ev_fw_panel);
  if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FAP_OK;
}

int fap_ctx_set_option(fap_ctx *ctx, const char *name, int64_t value) {
  if (!ctx || !name) return fail(FAP_E_ARG, "null argument");
  const std::string k(name);
  if (k == "fw_lookahead") {
    if (value < -1 || value > 1) return fail(FAP_E_ARG, "fw_lookahead must be -1 (auto), 0 or 1");
    ctx->opt_fw_lookahead = (int)value;
  } else if (k == "fw_tile") {
    if (value != 0 && value != 64 && value != 128)
      return fail(FAP_E_ARG, "fw_tile must be 0, 64 or 128");
    ctx->opt_fw_tile = (int)value;
  } else if (k == "sssp_threads") {
    if (value != 0 && value != 256 && value != 512 && value != 1024)
      return fail(FAP_E_ARG, "sssp_threads must be 0 (auto), 256, 512 or 1024");
    ctx->opt_sssp_threads = (int)value;
  } else if (k == "sssp_delta") {
    if (value < 0 || value > INT32_MAX / 4) return fail(FAP_E_ARG, "sssp_delta out of range");
    ctx->opt_sssp_delta = (int)value;
  } else if (k == "sssp_group") {
    if (value != 0 && value != 1 && value != 4 && value != 8)
      return fail(FAP_E_ARG, "sssp_group must be 0 (auto), 1, 4 or 8");
    ctx->opt_sssp_group = (int)value;
  } else if (k == "sssp_bin") {
    if (value < 0 || value > 3) return fail(FAP_E_ARG, "sssp_bin must be 0..3");
    ctx->opt_sssp_bin = (int)value;
  } else if (k == "sssp_pack") {
    if (value < 0 || value > 2) return fail(FAP_E_ARG, "sssp_pack must be 0 (auto), 1 or 2");
    ctx->opt_sssp_pack = value == 0 ? 1 : (int)value;
  } else if (k == "sssp_wgs") {
    if (value < 0 || value > 8) return fail(FAP_E_ARG, "sssp_wgs must be 0 (auto) .. 8");
    ctx->opt_sssp_wgs = (int)value;
  } else if (k == "sssp_arcs") {
    if (value != 0 && value != 1 && value != 2 && value != 4)
      return fail(FAP_E_ARG, "sssp_arcs must be 0 (auto), 1, 2 or 4");
    ctx->opt_sssp_arcs = value == 0 ? 2 : (int)value;
  } else if (k == "sssp_index") {
    if (value < 0 || value > 1) return fail(FAP_E_ARG, "sssp_index must be 0 or 1");
    ctx->opt_sssp_index = (int)value;
  } else if (k == "bidx_tile") {
    if (value != 0 && value != 64 && value != 128)
      return fail(FAP_E_ARG, "bidx_tile must be 0 (auto), 64 or 128");
    ctx->opt_bidx_tile = (int)value;
  } else if (k == "bidx_delta") {
    if (value < 0 || value > FAP_INF / 4) return fail(FAP_E_ARG, "bidx_delta out of range");
    ctx->opt_bidx_delta = (int)value;
  } else if (k == "batch_wait_us") {
    if (value < 0 || value > 600000000LL)
      return fail(FAP_E_ARG, "batch_wait_us must be 0 (auto) .. 600000000");
    ctx->opt_batch_wait_us = value;
  } else if (k == "batch_grid") {
    if (value < 0 || value > 64) return fail(FAP_E_ARG, "batch_grid must be 0 (auto) .. 64");
    ctx->opt_batch_grid = (int)value;
  } else {
    return fail(FAP_E_ARG, "unknown option '" + k + "'");
  }
  return FAP_OK;
}

int fap_ctx_get_option(fap_ctx *ctx, const char *name, int64_t *value) {
  if (!ctx || !name || !value) return fail(FAP_E_ARG, "null argument");
  const std::string k(name);
  if (k == "fw_lookahead") *value = ctx->opt_fw_lookahead;
  else if (k == "fw_tile") *value = ctx->opt_fw_tile;
  else if (k == "sssp_threads") *value = ctx->opt_sssp_threads;
  else if (k == "sssp_delta") *value = ctx->opt_sssp_delta;
  else if (k == "batch_grid") *value = ctx->opt_batch_grid;
  else if (k == "sssp_group") *value = ctx->opt_sssp_group;
  else if (k == "sssp_bin") *value = ctx->opt_sssp_bin;
  else if (k == "sssp_pack") *value = ctx->opt_sssp_pack;
  else if (k == "sssp_wgs") *value = ctx->opt_sssp_wgs;
  else if (k == "sssp_arcs") *value = ctx->opt_sssp_arcs;
  else if (k == "sssp_index") *value = ctx->opt_sssp_index;
  else if (k == "bidx_delta") *value = ctx->opt_bidx_delta;
  else if (k == "bidx_tile") *value = ctx->opt_bidx_tile;
  else if (k == "batch_wait_us") *value = ctx->opt_batch_wait_us;
  else return fail(FAP_E_ARG, "unknown option '" + k + "'");
  return FAP_OK;
}

int fap_ctx_stream(fap_ctx *ctx, void **stream) {
  if (!ctx || !stream) return fail(FAP_E_ARG, "null argument");
  *stream = (void *)ctx->stream;
  return FAP_OK;
}

int fap_ctx_sync(fap_ctx *ctx) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  FAP_HIP(hipStreamSynchronize(ctx->stream));
  return FAP_OK;
}

int fap_dev_alloc(fap_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return fail(FAP_E_ARG, "null argument");
  FAP_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 4);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return fail(FAP_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                 hipGetErrorString(e));
  }
  return FAP_OK;
}

int fap_dev_free(fap_ctx *ctx, void *dptr) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (dptr) FAP_HIP(hipFree(dptr));
  return FAP_OK;
}

int fap_memcpy_h2d(fap_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (!bytes) return FAP_OK;
  FAP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FAP_HIP(hipStreamSynchronize(ctx->stream));
  return FAP_OK;
}

int fap_memcpy_d2h(fap_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (!bytes) return FAP_OK;
  FAP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FAP_HIP(hipStreamSynchronize(ctx->stream));
  return FAP_OK;
}

int fap_memcpy_d2d(fap_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (!bytes) return FAP_OK;
  FAP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  FAP_HIP(hipStreamSynchronize(ctx->stream));
  return FAP_OK;
}

int fap_fill_i32(fap_ctx *ctx, int32_t *d, int64_t n, int32_t value, void *stream) {
  if (!ctx || (!d && n > 0)) return fail(FAP_E_ARG, "null argument");
  return launch_fill_i32(d, n, value, pick_stream(ctx, stream));
}

int fap_copy_i32(fap_ctx *ctx, int32_t *dst, const int32_t *src, int64_t n, void *stream) {
  if (!ctx || ((!dst || !src) && n > 0)) return fail(FAP_E_ARG, "null argument");
  if (n <= 0) return FAP_OK;
  const int vec = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
  const int64_t work = vec ? (n + 3) / 4 : n;
  const int64_t blocks = std::min<int64_t>((work + 255) / 256, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(copy_i32_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0,
                     pick_stream(ctx, stream), src, dst, n, vec);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int fap_checksum_i32(fap_ctx *ctx, const int32_t *d, int64_t n, uint64_t *sum,
                     uint64_t *weighted, void *stream) {
  if (!ctx || !sum || !weighted) return fail(FAP_E_ARG, "null argument");
  hipStream_t s = pick_stream(ctx, stream);
  FAP_TRY(ensure_ws(ctx, ctx->ws_misc, 64));
  unsigned long long *acc = (unsigned long long *)ctx->ws_misc.ptr;
  FAP_TRY(launch_fill_i32((int32_t *)acc, 4, 0, s));
  if (n > 0) {
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(checksum_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const uint32_t *)d, n, acc);
    FAP_CHECK_LAUNCH();
  }
  unsigned long long h[2];
  FAP_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  *sum = h[0];
  *weighted = h[1];
  return FAP_OK;
}

int fap_stream_create(fap_ctx *ctx, void **stream) {
  if (!ctx || !stream) return fail(FAP_E_ARG, "null argument");
  FAP_HIP(hipSetDevice(ctx->device));
  hipStream_t s;
  FAP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return FAP_OK;
}

int fap_stream_destroy(fap_ctx *ctx, void *stream) {
  (void)ctx;
  if (stream) FAP_HIP(hipStreamDestroy((hipStream_t)stream));
  return FAP_OK;
}

int fap_stream_sync(void *stream) {
  FAP_HIP(hipStreamSynchronize((hipStream_t)stream));
  return FAP_OK;
}

int fap_event_create(void **ev) {
  if (!ev) return fail(FAP_E_ARG, "null argument");
  hipEvent_t e;
  FAP_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return FAP_OK;
}

int fap_event_destroy(void *ev) {
  if (ev) FAP_HIP(hipEventDestroy((hipEvent_t)ev));
  return FAP_OK;
}

int fap_event_record(void *ev, void *stream) {
  FAP_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return FAP_OK;
}

int fap_event_elapsed(void *start, void *stop, float *ms) {
  if (!ms) return fail(FAP_E_ARG, "null argument");
  FAP_HIP(hipEventSynchronize((hipEvent_t)stop));
  FAP_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return FAP_OK;
}

// ---- resident graph ----

int fap_graph_upload(fap_ctx *ctx, int32_t V, int64_t E, const int32_t *row_offset,
                     const int32_t *col, const int32_t *w, fap_graph_dev **out) {
  if (!ctx || !out || !row_offset || (E > 0 && (!col || !w)))
    return fail(FAP_E_ARG, "null argument");
  if (V < 0 || E < 0 || E > INT32_MAX) return fail(FAP_E_ARG, "bad V/E");
  if (row_offset[0] != 0 || row_offset[V] != E)
    return fail(FAP_E_ARG, "row_offset must start at 0 and end at E");
  int32_t maxdeg = 0;
  for (int32_t v = 0; v < V; v++) {
    int32_t d = row_offset[v + 1] - row_offset[v];
    if (d < 0) return fail(FAP_E_ARG, "row_offset not monotone at " + std::to_string(v));
    maxdeg = std::max(maxdeg, d);
  }
  std::vector<int2> packed((size_t)E);
  int64_t wsum = 0;
  for (int64_t e = 0; e < E; e++) {
    if (col[e] < 0 || col[e] >= V)
      return fail(FAP_E_ARG, "col[" + std::to_string(e) + "] out of range");
    if (w[e] < 0 || w[e] >= FAP_INF)
      return fail(FAP_E_ARG, "weight[" + std::to_string(e) + "] outside [0, INF)");
    packed[e] = make_int2(col[e], w[e]);
    wsum += w[e];
  }
  FAP_HIP(hipSetDevice(ctx->device));
  fap_graph_dev *g = new fap_graph_dev();
  g->V = V;
  g->E = E;
  g->max_degree = maxdeg;
  g->weight_sum = wsum;
  hipError_t e1 = hipMalloc(&g->row_offset, sizeof(int32_t) * (V + 1));
  hipError_t e2 = hipMalloc(&g->edges, sizeof(int2) * (E > 0 ? E : 1));
  if (e1 != hipSuccess || e2 != hipSuccess) {
    if (g->row_offset) (void)hipFree(g->row_offset);
    if (g->edges) (void)hipFree(g->edges);
    delete g;
    (void)hipGetLastError();
    return fail(FAP_E_NOMEM, "graph upload: device allocation failed");
  }
  hipError_t e = hipMemcpyAsync(g->row_offset, row_offset, sizeof(int32_t) * (V + 1),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && E > 0)
    e = hipMemcpyAsync(g->edges, packed.data(), sizeof(int2) * E, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(g->row_offset);
    (void)hipFree(g->edges);
    delete g;
    return hip_fail(e, "graph upload copy", __FILE__, __LINE__);
  }
  *out = g;
  return FAP_OK;
}

int fap_graph_set_partition(fap_ctx *ctx, fap_graph_dev *g, const int32_t *graph_id,
                            const int32_t *st2ed, const int32_t *ed2st) {
  if (!ctx || !g || !graph_id || !st2ed || !ed2st) return fail(FAP_E_ARG, "null argument");
  const int32_t V = g->V;
  for (int32_t i = 0; i < V; i++) {
    if (st2ed[i] < 0 || st2ed[i] >= V || ed2st[i] < 0 || ed2st[i] >= V)
      return fail(FAP_E_ARG, "st2ed/ed2st entry out of range at " + std::to_string(i));
    if (ed2st[st2ed[i]] != i) return fail(FAP_E_ARG, "ed2st is not the inverse of st2ed");
  }
  FAP_HIP(hipSetDevice(ctx->device));
  int32_t **slots[3] = {&g->graph_id, &g->st2ed, &g->ed2st};
  const int32_t *src[3] = {graph_id, st2ed, ed2st};
  for (int k = 0; k < 3; k++) {
    if (!*slots[k]) FAP_HIP(hipMalloc(slots[k], sizeof(int32_t) * (V > 0 ? V : 1)));
    if (V > 0)
      FAP_HIP(hipMemcpyAsync(*slots[k], src[k], sizeof(int32_t) * V, hipMemcpyHostToDevice,
                             ctx->stream));
  }
  FAP_HIP(hipStreamSynchronize(ctx->stream));
  g->h_st2ed.assign(st2ed, st2ed + V);
  g->h_graph_id.assign(graph_id, graph_id + V);
  bidx_release(g);  // an index belongs to the partition it was built for
  return FAP_OK;
}

int fap_graph_free(fap_ctx *ctx, fap_graph_dev *g) {
  if (!g) return FAP_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  bidx_release(g);
  int32_t *bufs[] = {g->row_offset, g->graph_id, g->st2ed, g->ed2st, g->in_offset};
  for (int32_t *b : bufs)
    if (b) (void)hipFree(b);
  if (g->edges) (void)hipFree(g->edges);
  if (g->in_arcs) (void)hipFree(g->in_arcs);
  delete g;
  return FAP_OK;
}

int fap_graph_info(const fap_graph_dev *g, int32_t *V, int64_t *E) {
  if (!g) return fail(FAP_E_ARG, "graph is null");
  if (V) *V = g->V;
  if (E) *E = g->E;
  return FAP_OK;
}

}  // extern "C"
