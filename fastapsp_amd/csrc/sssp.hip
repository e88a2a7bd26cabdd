// sssp.hip -- multi-source SSSP (stage 1) by frontier Bellman-Ford.
//
// Replaces handle_boundry_path (fap/kernel/batch_sssp.h:254-277), whose CPU
// branch runs one binary-heap Dijkstra per boundary vertex (:79-114, OpenMP
// over sources :232-252) and whose HIP branch runs the sources one after
// another with a host round trip per relaxation round
// (fap/kernel/hip/batch_sssp.cpp:91-158).  Shortest distances are unique, so
// any exact method gives the reference's rows bit for bit.
//
// Here all sources of a call advance together.  The frontier is a queue of
// (source, vertex) items, item = s * V + v (uint32; sources are processed in
// chunks so that S * V < 2^32 and the per-chunk state fits the budget).
// Round r: each item (s, u) reads dist[s][u] and relaxes u's out-arcs with
//   atomicMin(dist[s][v], dist[s][u] + w)     (only when a plain read shows
//                                              the candidate is smaller)
// and an improved (s, v) is queued for round r+1 unless already queued for it
// (atomicMax on a per-(s, v) round stamp).  Queue appends are compacted per
// wavefront: ballot of the lanes that push, one atomicAdd per wave, lane slot
// = mbcnt of the ballot.  Arcs are read as packed {col, w} pairs.
//
// The host never waits per round: rounds are launched back to back, each
// kernel reads its queue length from device memory, and every POLL rounds an
// asynchronous copy of the next queue length is checked one poll late; the
// rounds after convergence are empty launches.  Bellman-Ford with
// non-negative weights settles within V rounds; more rounds than V + 1 is
// reported as FAP_E_DEVICE.
#include <algorithm>
#include <vector>

#include "common.h"

namespace fapk {

constexpr int SSSP_THREADS = 256;

__global__ __launch_bounds__(SSSP_THREADS) void sssp_init_kernel(
    int32_t *dist, int64_t ld, int V, int S, const int32_t *sources, int32_t *stamp,
    uint32_t *q0, uint32_t *counts) {
  const int64_t total = (int64_t)S * V;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int s = (int)(i / V);
    const int v = (int)(i - (int64_t)s * V);
    dist[(int64_t)s * ld + v] = (v == sources[s]) ? 0 : 0x3FFFFFFF;
    stamp[i] = 0;
  }
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += stride)
    q0[s] = (uint32_t)(s * V + sources[s]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts[0] = (uint32_t)S;
    counts[1] = 0;
    counts[2] = 0;
  }
}

__device__ __forceinline__ unsigned lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__global__ __launch_bounds__(SSSP_THREADS) void sssp_round_kernel(
    const int32_t *__restrict__ row_offset, const int2 *__restrict__ edges, int32_t *dist,
    int64_t ld, int V, const uint32_t *__restrict__ qcur, uint32_t *__restrict__ qnext,
    uint32_t *counts, int32_t *stamp, int r) {
  const uint32_t n = counts[r % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) counts[(r + 2) % 3] = 0;
  uint32_t *next_count = counts + (r + 1) % 3;
  const int stamp_val = r + 1;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t idx = base + threadIdx.x;
    const bool valid = idx < n;
    int s = 0, u = 0, du = 0, e = 0, e_end = 0;
    if (valid) {
      const uint32_t item = qcur[idx];
      s = (int)(item / (uint32_t)V);
      u = (int)(item - (uint32_t)s * (uint32_t)V);
      du = dist[(int64_t)s * ld + u];
      e = row_offset[u];
      e_end = row_offset[u + 1];
    }
    while (__any(valid && e < e_end)) {
      bool push = false;
      uint32_t pitem = 0;
      if (valid && e < e_end) {
        const int2 ed = edges[e];
        e++;
        const int nd = du + ed.y;
        int32_t *dp = dist + (int64_t)s * ld + ed.x;
        if (nd < *dp) {
          const int old = atomicMin(dp, nd);
          if (nd < old) {
            const uint32_t key = (uint32_t)s * (uint32_t)V + (uint32_t)ed.x;
            if (atomicMax(stamp + key, stamp_val) < stamp_val) {
              push = true;
              pitem = key;
            }
          }
        }
      }
      const unsigned long long mask = __ballot(push);
      if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t wbase = 0;
        if (lane == leader) wbase = atomicAdd(next_count, (uint32_t)__popcll(mask));
        wbase = __shfl(wbase, leader);
        if (push) qnext[wbase + lane_prefix(mask)] = pitem;
      }
    }
  }
}

}  // namespace fapk

namespace fapi {

static constexpr int kPoll = 16;

// Runs one chunk of sources (S * V < 2^32, state already sized).
static int sssp_chunk(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int S,
                      int32_t *d_out, int64_t ld, hipStream_t s, uint32_t *q0, uint32_t *q1,
                      int32_t *stamp, uint32_t *counts, hipEvent_t ev[2]) {
  const int V = g->V;
  const int64_t total = (int64_t)S * V;
  int64_t blocks = std::min<int64_t>((total + fapk::SSSP_THREADS - 1) / fapk::SSSP_THREADS,
                                     (int64_t)ctx->num_cus * 16);
  blocks = std::max<int64_t>(blocks, 1);
  hipLaunchKernelGGL(fapk::sssp_init_kernel, dim3((unsigned)blocks), dim3(fapk::SSSP_THREADS),
                     0, s, d_out, ld, V, S, d_sources, stamp, q0, counts);
  FAP_CHECK_LAUNCH();
  const unsigned round_blocks = (unsigned)std::max(1, ctx->num_cus * 8);
  uint32_t *slot = ctx->host_pinned;  // two poll slots
  const int max_rounds = V + 2;
  int pending = -1;  // poll index whose event is outstanding
  for (int r = 0;; r++) {
    if (r > max_rounds)
      return fail(FAP_E_DEVICE, "sssp: no convergence within V + 1 rounds");
    uint32_t *qc = (r & 1) ? q1 : q0;
    uint32_t *qn = (r & 1) ? q0 : q1;
    hipLaunchKernelGGL(fapk::sssp_round_kernel, dim3(round_blocks), dim3(fapk::SSSP_THREADS),
                       0, s, g->row_offset, g->edges, d_out, ld, V, qc, qn, counts, stamp, r);
    FAP_CHECK_LAUNCH();
    if ((r + 1) % kPoll == 0) {
      const int p = ((r + 1) / kPoll) & 1;
      FAP_HIP(hipMemcpyAsync(slot + p, counts + (r + 1) % 3, sizeof(uint32_t),
                             hipMemcpyDeviceToHost, s));
      FAP_HIP(hipEventRecord(ev[p], s));
      if (pending >= 0) {
        FAP_HIP(hipEventSynchronize(ev[pending]));
        if (slot[pending] == 0) break;
      }
      pending = p;
    }
  }
  return FAP_OK;
}

int run_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
             int32_t *d_out, int64_t ld_out, hipStream_t s) {
  if (n_src <= 0) return FAP_OK;
  const int V = g->V;
  if (V <= 0) return fail(FAP_E_ARG, "sssp: empty graph");
  // chunk so that S*V < 2^32 and the queue+stamp state stays within ~4 GiB
  const int64_t by_index = (int64_t)(UINT32_MAX - 1) / V;
  const int64_t by_mem = std::max<int64_t>(1, ((int64_t)4 << 30) / (12 * (int64_t)V));
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>({by_index, by_mem, n_src}));
  const size_t cells = (size_t)chunk * V;
  FAP_TRY(ensure_ws(ctx, ctx->ws_queue, 2 * cells * sizeof(uint32_t)));
  FAP_TRY(ensure_ws(ctx, ctx->ws_stamp, cells * sizeof(int32_t)));
  FAP_TRY(ensure_ws(ctx, ctx->ws_small, 64));
  uint32_t *q0 = (uint32_t *)ctx->ws_queue.ptr;
  uint32_t *q1 = q0 + cells;
  int32_t *stamp = (int32_t *)ctx->ws_stamp.ptr;
  uint32_t *counts = (uint32_t *)ctx->ws_small.ptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  FAP_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  hipError_t he = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (he != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return hip_fail(he, "hipEventCreate", __FILE__, __LINE__);
  }
  int rc = FAP_OK;
  for (int32_t s0 = 0; s0 < n_src && rc == FAP_OK; s0 += chunk) {
    const int S = std::min<int32_t>(chunk, n_src - s0);
    rc = sssp_chunk(ctx, g, d_sources + s0, S, d_out + (int64_t)s0 * ld_out, ld_out, s, q0, q1,
                    stamp, counts, ev);
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}

}  // namespace fapi

using namespace fapi;

extern "C" int fap_sssp_multi(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources,
                              int32_t n_src, int32_t *d_out, int64_t ld_out, void *stream) {
  if (!ctx || !g) return fail(FAP_E_ARG, "null ctx/graph");
  if (n_src < 0) return fail(FAP_E_ARG, "sssp: negative source count");
  if (n_src == 0) return FAP_OK;
  if (!d_sources || !d_out) return fail(FAP_E_ARG, "sssp: null pointer");
  if (ld_out < g->V) return fail(FAP_E_ARG, "sssp: ld_out < V");
  hipStream_t s = pick_stream(ctx, stream);
  // validate the sources on the host: an out-of-range id would index outside
  // the output rows
  std::vector<int32_t> h(n_src);
  FAP_HIP(hipMemcpyAsync(h.data(), d_sources, sizeof(int32_t) * n_src, hipMemcpyDeviceToHost, s));
  FAP_HIP(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n_src; i++)
    if (h[i] < 0 || h[i] >= g->V)
      return fail(FAP_E_ARG, "sssp: source " + std::to_string(i) + " out of range");
  return run_sssp(ctx, g, d_sources, n_src, d_out, ld_out, s);
}
