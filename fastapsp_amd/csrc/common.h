// common.h -- internal helpers shared by the HIP and host translation units.
#ifndef FAP_COMMON_H_
#define FAP_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "fap_dev.h"
#include "fap_tune.h"

namespace fapi {

constexpr int32_t kInf = FAP_INF;

// Thread-local last error; every C entry point returns a status and leaves
// the message here.
void set_error(const std::string &msg);
void clear_error();
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define FAP_HIP(call)                                                  \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) return fapi::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define FAP_CHECK_LAUNCH()                                             \
  do {                                                                 \
    hipError_t e_ = hipGetLastError();                                 \
    if (e_ != hipSuccess) return fapi::hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
  } while (0)

#define FAP_TRY(call)                                                  \
  do {                                                                 \
    int rc_ = (call);                                                  \
    if (rc_ != FAP_OK) return rc_;                                     \
  } while (0)

// Device buffer grown on demand (never shrinks); owned by the context.
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace fapi

struct fap_bidx;  // boundary index of a partitioned graph (bindex.hip)

struct fap_graph_dev {
  int32_t V = 0;
  int64_t E = 0;
  int32_t *row_offset = nullptr;  // V+1
  int2 *edges = nullptr;          // E x {col, w}
  int32_t *graph_id = nullptr;    // V (after set_partition)
  int32_t *st2ed = nullptr;       // V
  int32_t *ed2st = nullptr;       // V
  int32_t max_degree = 0;
  int64_t weight_sum = 0;
  int32_t delta = 0;  // SSSP near-far band width (0 = derived from the mean weight)
  int32_t *in_offset = nullptr;  // transposed CSR (V+1), built on first path request
  int2 *in_arcs = nullptr;       // E x {source, w}
  std::vector<int32_t> h_st2ed;     // host copy of st2ed (source lists of batched solves)
  std::vector<int32_t> h_graph_id;  // host copy of graph_id (part range checks)
  fap_bidx *bidx = nullptr;         // built on first batched part solve (bindex.hip)
  int bidx_state = 0;               // 0 not tried, 1 built, -1 not applicable
};

struct fap_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int xcc_mask = 0;  // XCDs the device's workgroups run on (bits of XCC ids); 0 = not probed
  size_t lds_per_block = 0;
  std::mutex mu;  // serialises workspace growth
  // named workspaces (device)
  fapi::Workspace ws_sub;     // dense part matrix
  fapi::Workspace ws_queue;   // SSSP frontier queues
  fapi::Workspace ws_stamp;   // SSSP enqueue stamps
  fapi::Workspace ws_small;   // counters etc.
  fapi::Workspace ws_misc;
  fapi::Workspace ws_batch;   // batched-closure descriptors and counters
  fapi::Workspace ws_dense;   // dense-build descriptors
  fapi::Workspace ws_srcs;    // batched-solve sources and row pointers
  fapi::Workspace ws_bidx;    // boundary distances of a batched solve's sources
  fapi::Workspace ws_bbf;     // boundary iterations: tile lists and changed flags
  uint32_t *host_pinned = nullptr;  // small pinned host buffer for counters
  void *host_bbf = nullptr;         // pinned staging of the boundary iterations
  size_t host_bbf_bytes = 0;
  hipEvent_t ev_poll = nullptr;
  // second stream and events for the Floyd-Warshall look-ahead
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fw_la = nullptr;
  hipEvent_t ev_fw_panel = nullptr;
  // tuning options (fap_ctx_set_option); they change speed, never results.
  // "auto" values defer to the tune cache, then to the defaults.
  int opt_fw_lookahead = -1;     // run round r+1's phases 1-2 beside round r's phase 3; -1 auto
  int opt_fw_tile = 0;           // FW block size 64 / 128; 0 auto
  int opt_sssp_threads = 0;      // threads per SSSP workgroup: 256, 512, 1024; 0 auto (by count)
  int opt_sssp_delta = 0;        // near-far band; 0 auto (graph's own, cache, 8 x mean weight)
  int opt_sssp_group = 0;        // sources per SSSP workgroup pass: 1, 4, 8; 0 = 1
  int opt_batch_grid = 0;        // batched-closure workgroups per CU per launch; 0 auto
  int opt_sssp_bin = 3;          // reorder SSSP queues by vertex block from 2NT / NT / NT/2 entries (1/2/3); 0 off
  int opt_sssp_pack = 1;         // SSSP state: 64-bit words (1), in-row 32-bit words (2)
  int opt_sssp_wgs = 0;          // SSSP workgroups per CU: 1..8; 0 = 1024 threads per CU
  int opt_sssp_arcs = 2;         // arcs per lane per SSSP relaxation step: 1, 2, 4
  int opt_sssp_index = 1;        // batched part solves: stage 1 through the boundary index (1) or SSSP (0)
  int opt_bidx_tile = 0;         // boundary iterations' tile columns: 64 or 128; 0 = 128
  int opt_bidx_delta = 0;        // distance band of the boundary iterations; 0 = mean arc weight x sqrt(V / K)
  int64_t opt_batch_wait_us = 0;  // single-launch batched closure: longest wait of a task; 0 = 2 s
  fap_tune_cache *tune = nullptr;  // not owned
  // pinned host staging for asynchronous descriptor uploads; free for reuse
  // once ev_stage (recorded after the last upload from it) has completed
  void *host_stage = nullptr;
  size_t host_stage_bytes = 0;
  hipEvent_t ev_stage = nullptr;
  bool stage_pending = false;
  // recorded after a batched part solve's SSSP launch, so that the next
  // group's SSSP (on the peer's stream) starts after it, not beside it
  hipEvent_t ev_sssp = nullptr;
  // second context (own stream and workspaces) for two-stream pipelines of
  // batched part solves; created on first use, owned
  fap_ctx *peer = nullptr;
};

namespace fapi {
// Ensure ws holds >= bytes (device memory of ctx's device). Synchronizes the
// device before reallocating.
int ensure_ws(fap_ctx *ctx, Workspace &ws, size_t bytes);
inline hipStream_t pick_stream(fap_ctx *ctx, void *stream) {
  return stream ? (hipStream_t)stream : ctx->stream;
}
// Effective tuning settings: the context option when set, else the tune
// cache entry for the bucket, else the default (tune_cache.cpp).
bool tune_lookup(const fap_ctx *ctx, const char *kind, int64_t key, int32_t *params);
fap_tune_cache *tune_env_cache();
void fw_settings(const fap_ctx *ctx, int32_t n, int *tile, bool *lookahead);
void sssp_settings(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src, int *threads,
                   int *delta, int *group);
int batch_grid_per_cu(const fap_ctx *ctx, int64_t batch_key);
// Pinned host staging of at least `bytes`, free to overwrite: waits for the
// previous upload from it, grows it when needed (ctx.hip).
int acquire_host_stage(fap_ctx *ctx, size_t bytes, void **out);
// Marks the staging in use by an upload just enqueued on stream s.
int release_host_stage(fap_ctx *ctx, hipStream_t s);
// Device-side helpers implemented in ctx.hip.
int launch_fill_i32(int32_t *d, int64_t n, int32_t value, hipStream_t s);
int launch_fill_u32(uint32_t *d, int64_t n, uint32_t value, hipStream_t s);
// Boundary index (bindex.hip): build (synchronous; leaves g->bidx null when
// the graph does not qualify), stage-1 rows through it (waits on s), release.
int bidx_build(fap_ctx *ctx, fap_graph_dev *g);
int bidx_rows(fap_ctx *ctx, const fap_graph_dev *g, int32_t count, const int32_t *qidx,
              const int32_t *rowoff, int32_t nsrc, int32_t *const *d_rowptr, hipStream_t s);
void bidx_release(fap_graph_dev *g);
// Index of the part starting at st2ed offset `start` with n vertices and b
// boundary ones, or -1 when the index does not hold that part as given.
int32_t bidx_part(const fap_graph_dev *g, int32_t start, int32_t n, int32_t b);
}  // namespace fapi

namespace fapk {
// Map a linear workgroup id to a work index so that each XCD (workgroups are
// dealt round-robin over the MI355X's 8 XCDs) takes a contiguous range of
// work, whose shared inputs then meet in that XCD's L2.  A bijection on
// [0, total): ids past the last multiple of 8 map to themselves.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
  const int per = total >> 3;
  if (bid >= (per << 3)) return bid;
  return (bid & 7) * per + (bid >> 3);
}
}  // namespace fapk

#endif  // FAP_COMMON_H_
