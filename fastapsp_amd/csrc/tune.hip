// tune.hip -- the autotuner's measurements (include/fap_tune.h).
//
// Each tuner times every candidate setting on the device with events on the
// context stream (one warm-up, then `reps` timed runs) and records the
// fastest in the context's cache under the workload's bucket.  Candidates
// only change launch shapes; every one computes the same distances.
#include <algorithm>
#include <string>
#include <vector>

#include "batch_desc.h"
#include "common.h"
#include "fap_tune.h"

namespace fapk {

// The dense benchmark's kind of matrix, from a hash of (i, j, seed): an arc
// with probability 1/64 and weight 1..1000, else INF; zero diagonal.
__global__ void tune_fill_kernel(int32_t *__restrict__ d, int32_t n, int64_t ld, uint32_t seed) {
  const int64_t cells = (int64_t)n * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < cells; x += stride) {
    const int64_t i = x / n, j = x - (x / n) * n;
    uint32_t h = (uint32_t)(x * 0x9E3779B97F4A7C15ull >> 32) ^ seed;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    d[i * ld + j] = (i == j) ? 0 : ((h & 63) == 0 ? (int32_t)(1 + (h >> 8) % 1000) : FAP_INF);
  }
}

// Issue-rate probe for the min-plus instruction mix: per update one
// v_add_u32, per two updates one v_min3_i32 folding both sums into the
// accumulator -- the inner loop of the min-plus tile without its LDS reads.
// Eight independent accumulators per lane; inline asm keeps the compiler
// from folding the chains.  Each thread writes only its own output word.
__global__ __launch_bounds__(256) void valu_mix_kernel(int32_t *out, int iters, int seed) {
  int a0 = threadIdx.x + seed, a1 = a0 ^ 1, a2 = a0 ^ 2, a3 = a0 ^ 3;
  int a4 = a0 ^ 4, a5 = a0 ^ 5, a6 = a0 ^ 6, a7 = a0 ^ 7;
  const int b = seed * 3 + 1, c = seed * 5 + 2;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int t0, t1;
#define FAP_MIX(acc, x, y)                                                           \
  asm volatile("v_add_u32 %0, %1, %2" : "=v"(t0) : "v"(x), "v"(b));                 \
  asm volatile("v_add_u32 %0, %1, %2" : "=v"(t1) : "v"(y), "v"(c));                 \
  asm volatile("v_min3_i32 %0, %0, %1, %2" : "+v"(acc) : "v"(t0), "v"(t1));
      FAP_MIX(a0, a1, a2) FAP_MIX(a1, a2, a3) FAP_MIX(a2, a3, a4) FAP_MIX(a3, a4, a5)
      FAP_MIX(a4, a5, a6) FAP_MIX(a5, a6, a7) FAP_MIX(a6, a7, a0) FAP_MIX(a7, a0, a1)
#undef FAP_MIX
    }
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

}  // namespace fapk

namespace fapi {

int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);
int fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *mats, const int32_t *n,
                   const int64_t *ld, int use_graph, fap_fw_plan **out);
int fw_plan_run(fap_fw_plan *p, hipStream_t s);
void fw_plan_destroy(fap_fw_plan *p);

namespace {

// Device scratch freed on scope exit.
struct Scratch {
  void *p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(FAP_E_NOMEM, "tuner scratch of " + std::to_string(bytes) + " bytes");
    }
    return FAP_OK;
  }
};

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int create() {
    FAP_HIP(hipEventCreate(&a));
    FAP_HIP(hipEventCreate(&b));
    return FAP_OK;
  }
};

int fill(int32_t *d, int32_t n, int64_t ld, uint32_t seed, hipStream_t s) {
  if (n <= 0) return FAP_OK;
  const int64_t cells = (int64_t)n * n;
  hipLaunchKernelGGL(fapk::tune_fill_kernel, dim3((unsigned)std::min<int64_t>((cells + 255) / 256, 4096)),
                     dim3(256), 0, s, d, n, ld, seed);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// Average device milliseconds of run() over reps runs after one warm-up;
// prep() runs before each, outside the timed region.
template <typename Prep, typename Run>
int time_runs(fap_ctx *ctx, Events &ev, int reps, Prep prep, Run run, double *ms) {
  hipStream_t s = ctx->stream;
  FAP_TRY(prep());
  FAP_TRY(run());
  double total = 0;
  for (int r = 0; r < reps; r++) {
    FAP_TRY(prep());
    FAP_HIP(hipEventRecord(ev.a, s));
    FAP_TRY(run());
    FAP_HIP(hipEventRecord(ev.b, s));
    FAP_HIP(hipEventSynchronize(ev.b));
    float t = 0;
    FAP_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
    total += t;
  }
  *ms = total / reps;
  return FAP_OK;
}

int need_cache(fap_ctx *ctx) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  if (!ctx->tune) return fail(FAP_E_STATE, "context has no tune cache (fap_ctx_set_tune_cache)");
  FAP_HIP(hipSetDevice(ctx->device));
  return FAP_OK;
}

}  // namespace
}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_tune_fw(fap_ctx *ctx, int32_t n, int32_t reps) {
  FAP_TRY(need_cache(ctx));
  if (n < 1 || n > 46340) return fail(FAP_E_ARG, "tune_fw: n outside 1..46340");
  reps = std::max(reps, 1);
  const int64_t ld = ((int64_t)n + 3) & ~(int64_t)3;
  const size_t bytes = (size_t)n * ld * sizeof(int32_t);
  Scratch base, work;
  FAP_TRY(base.alloc(bytes));
  FAP_TRY(work.alloc(bytes));
  Events ev;
  FAP_TRY(ev.create());
  hipStream_t s = ctx->stream;
  FAP_TRY(fill((int32_t *)base.p, n, ld, 12345u, s));
  const int saved_tile = ctx->opt_fw_tile, saved_la = ctx->opt_fw_lookahead;
  double best = -1;
  int32_t bp[2] = {64, 0};
  int rc = FAP_OK;
  for (int tile : {64, 128}) {
    for (int la : {0, 1}) {
      ctx->opt_fw_tile = tile;
      ctx->opt_fw_lookahead = la;
      double ms = 0;
      rc = time_runs(
          ctx, ev, reps,
          [&]() -> int {
            FAP_HIP(hipMemcpyAsync(work.p, base.p, bytes, hipMemcpyDeviceToDevice, s));
            return FAP_OK;
          },
          [&] { return launch_fw(ctx, (int32_t *)work.p, n, ld, s); }, &ms);
      if (rc != FAP_OK) break;
      if (best < 0 || ms < best) {
        best = ms;
        bp[0] = tile;
        bp[1] = la;
      }
    }
    if (rc != FAP_OK) break;
  }
  ctx->opt_fw_tile = saved_tile;
  ctx->opt_fw_lookahead = saved_la;
  FAP_TRY(rc);
  return fap_tune_cache_set(ctx->tune, "fw", fap_tune_key_fw(n), bp, 2, best);
}

int fap_tune_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *h_sources, int32_t n_src,
                  int32_t reps) {
  FAP_TRY(need_cache(ctx));
  if (!g || !h_sources || n_src < 1) return fail(FAP_E_ARG, "tune_sssp: bad arguments");
  for (int32_t i = 0; i < n_src; i++)
    if (h_sources[i] < 0 || h_sources[i] >= g->V)
      return fail(FAP_E_ARG, "tune_sssp: source out of range");
  reps = std::max(reps, 1);
  Scratch src, out;
  FAP_TRY(src.alloc(sizeof(int32_t) * n_src));
  FAP_TRY(out.alloc(sizeof(int32_t) * (size_t)n_src * g->V));
  Events ev;
  FAP_TRY(ev.create());
  hipStream_t s = ctx->stream;
  FAP_HIP(hipMemcpyAsync(src.p, h_sources, sizeof(int32_t) * n_src, hipMemcpyHostToDevice, s));
  FAP_HIP(hipStreamSynchronize(s));
  const double mean = g->E ? (double)g->weight_sum / (double)g->E : 1.0;
  const int saved_t = ctx->opt_sssp_threads, saved_d = ctx->opt_sssp_delta;
  double best = -1;
  int32_t bp[2] = {1024, 8};
  int rc = FAP_OK;
  for (int threads : {256, 512, 1024}) {
    for (int mult : {2, 4, 8, 16, 32}) {
      ctx->opt_sssp_threads = threads;
      ctx->opt_sssp_delta = (int)std::max(1.0, std::min(1.0e6, mult * mean));
      double ms = 0;
      rc = time_runs(
          ctx, ev, reps, [] { return FAP_OK; },
          [&]() -> int {
            return run_sssp_rows(ctx, g, (const int32_t *)src.p, n_src, (int32_t *)out.p, g->V,
                                 nullptr, s);
          },
          &ms);
      if (rc != FAP_OK) break;
      if (best < 0 || ms < best) {
        best = ms;
        bp[0] = threads;
        bp[1] = mult;
      }
    }
    if (rc != FAP_OK) break;
  }
  ctx->opt_sssp_threads = saved_t;
  ctx->opt_sssp_delta = saved_d;
  FAP_TRY(rc);
  return fap_tune_cache_set(ctx->tune, "sssp", fap_tune_key_sssp(g->V, g->E, n_src), bp, 2, best);
}

int fap_tune_batch(fap_ctx *ctx, int32_t count, const int32_t *n, int32_t reps) {
  FAP_TRY(need_cache(ctx));
  if (count < 1 || !n) return fail(FAP_E_ARG, "tune_batch: bad arguments");
  reps = std::max(reps, 1);
  std::vector<int64_t> ld(count), off(count);
  size_t words = 0;
  for (int32_t j = 0; j < count; j++) {
    if (n[j] < 0 || n[j] > 46340) return fail(FAP_E_ARG, "tune_batch: bad size");
    ld[j] = ((int64_t)n[j] + 3) & ~(int64_t)3;
    off[j] = (int64_t)words;
    words += (size_t)n[j] * ld[j];
  }
  Scratch base, work;
  FAP_TRY(base.alloc(words * 4));
  FAP_TRY(work.alloc(words * 4));
  Events ev;
  FAP_TRY(ev.create());
  hipStream_t s = ctx->stream;
  std::vector<int32_t *> mats(count);
  for (int32_t j = 0; j < count; j++) {
    FAP_TRY(fill((int32_t *)base.p + off[j], n[j], ld[j], 777u + j, s));
    mats[j] = (int32_t *)work.p + off[j];
  }
  const int saved = ctx->opt_batch_grid;
  double best = -1;
  int32_t bp[1] = {64};
  int rc = FAP_OK;
  for (int per_cu : {4, 8, 16, 32, 64}) {
    ctx->opt_batch_grid = per_cu;
    fap_fw_plan *plan = nullptr;
    rc = fw_plan_create(ctx, count, mats.data(), n, ld.data(), 0, &plan);
    if (rc != FAP_OK) break;
    double ms = 0;
    rc = time_runs(
        ctx, ev, reps,
        [&]() -> int {
          FAP_HIP(hipMemcpyAsync(work.p, base.p, words * 4, hipMemcpyDeviceToDevice, s));
          return FAP_OK;
        },
        [&] { return fw_plan_run(plan, s); }, &ms);
    fw_plan_destroy(plan);
    if (rc != FAP_OK) break;
    if (best < 0 || ms < best) {
      best = ms;
      bp[0] = per_cu;
    }
  }
  ctx->opt_batch_grid = saved;
  FAP_TRY(rc);
  return fap_tune_cache_set(ctx->tune, "batch", fap_tune_key_batch(count, n), bp, 1, best);
}

int fap_probe_minplus_ceiling(fap_ctx *ctx, double *tops) {
  if (!ctx || !tops) return fail(FAP_E_ARG, "null argument");
  // 8 workgroups of 256 threads per CU (8 waves per SIMD), 16 updates per
  // unrolled step and lane: 64 steps x iters, ~15 ms per launch; one
  // warm-up launch, then the fastest of three
  const int blocks = ctx->num_cus * 8, iters = 8000;
  Scratch out;
  FAP_TRY(out.alloc(sizeof(int32_t) * (size_t)blocks * 256));
  hipStream_t s = ctx->stream;
  Events ev;
  FAP_TRY(ev.create());
  hipLaunchKernelGGL(fapk::valu_mix_kernel, dim3(blocks), dim3(256), 0, s, (int32_t *)out.p,
                     iters, 1);
  FAP_CHECK_LAUNCH();
  float best = 0;
  for (int rep = 0; rep < 3; rep++) {
    FAP_HIP(hipEventRecord(ev.a, s));
    hipLaunchKernelGGL(fapk::valu_mix_kernel, dim3(blocks), dim3(256), 0, s, (int32_t *)out.p,
                       iters, 2 + rep);
    FAP_CHECK_LAUNCH();
    FAP_HIP(hipEventRecord(ev.b, s));
    FAP_HIP(hipEventSynchronize(ev.b));
    float ms = 0;
    FAP_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (rep == 0 || ms < best) best = ms;
  }
  const double updates = (double)blocks * 256 * iters * 8 * 16;
  *tops = 2.0 * updates / (best * 1e-3) / 1e12;
  return FAP_OK;
}

}  // extern "C"
