// serve.hip -- request loop with continuous batching (include/fap_serve.h).
//
// No counterpart in the reference: it replaces the multi-task MPI driver
// (fap/fap_more_task_mpi.cpp:34-91) on one device and answers point
// queries (getDistanceP2P, a stub at fap/fap.cpp:263-265).
//
// One host thread runs the loop:
//   1. completions: a lane's wave is done when its event is; query answers
//      are copied out of the lane's pinned staging and the query rows freed;
//   2. admissions: every free lane takes a wave from the queue -- requests
//      of the head's kind, in arrival order, while the arena, the staging
//      and the wave limits allow -- stages it and launches it;
//   3. waiting: on the queue when no lane is busy, else a short poll.
// Device memory for results comes from one arena (first fit, coalescing),
// so admission is by footprint and no hipMalloc/hipFree runs per request.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_desc.h"
#include "common.h"
#include "fap_serve.h"
#include "host.h"

namespace fapk {

// ans[q] = row of pair q's source, at pair q's destination column
__global__ void answers_kernel(int32_t *const *__restrict__ rowptr,
                               const int32_t *__restrict__ qrow,
                               const int32_t *__restrict__ qdst, int32_t *__restrict__ ans,
                               int nq) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nq) ans[q] = rowptr[qrow[q]][qdst[q]];
}

// Point queries from the index: the cell (s, t) of s's part block, as the
// reference pipeline would produce it (fap/fap.cpp:92-181).  pinfo[p] =
// {part offset in st2ed order, size, boundary count, first boundary row in
// B}; M holds each part's closed matrix (row stride = size rounded up to 4)
// at moff[p]; B the SSSP rows of all boundary vertices.
__global__ void indexed_answers_kernel(const int32_t *__restrict__ gid,
                                       const int32_t *__restrict__ ed2st,
                                       const int4 *__restrict__ pinfo,
                                       const int64_t *__restrict__ moff,
                                       const int32_t *__restrict__ M, const int32_t *__restrict__ B,
                                       int64_t V, const int32_t *__restrict__ qsrc,
                                       const int32_t *__restrict__ qdst, int32_t *__restrict__ ans,
                                       int nq) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int s = qsrc[q], t = qdst[q];
  const int p = gid[s];
  const int4 pi = pinfo[p];
  const int i = ed2st[s] - pi.x;
  const int64_t ld = ((int64_t)pi.y + 3) & ~(int64_t)3;
  const int32_t *row = M + moff[p] + (int64_t)i * ld;
  int r;
  if (gid[t] == p) {
    r = row[ed2st[t] - pi.x];
  } else if (i < pi.z) {
    r = B[((int64_t)pi.w + i) * V + t];
  } else {
    r = 0x3FFFFFFF;
    for (int k = 0; k < pi.z; k++) r = min(r, row[k] + B[((int64_t)pi.w + k) * V + t]);
  }
  ans[q] = r;
}

// fap_checksum_i32 of an n x n matrix stored with row stride ld, indexed as
// if packed (cell (i, j) has position i * n + j).
__global__ void checksum2d_kernel(const uint32_t *__restrict__ d, int64_t ld, int32_t n,
                                  unsigned long long *out) {
  unsigned long long s = 0, ws = 0;
  const int64_t cells = (int64_t)n * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < cells; x += stride) {
    const int64_t i = x / n, j = x - (x / n) * n;
    const unsigned long long v = d[i * ld + j];
    s += v;
    ws += v * (unsigned long long)(x + 1);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    ws += __shfl_down(ws, off);
  }
  __shared__ unsigned long long ps[4], pw[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    ps[wv] = s;
    pw[wv] = ws;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(out, ps[0] + ps[1] + ps[2] + ps[3]);
    atomicAdd(out + 1, pw[0] + pw[1] + pw[2] + pw[3]);
  }
}

}  // namespace fapk

namespace fapi {

int launch_fw(fap_ctx *ctx, int32_t *d, int32_t n, int64_t ld, hipStream_t s);

// graphs at least this large are closed by the dense blocked FW on their own
constexpr int32_t DENSE_PATH_N = 3072;
namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation.  Offsets are 256-byte
// aligned; freed blocks merge with free neighbours.
class Arena {
 public:
  void init(size_t size) {
    size_ = size;
    used_ = peak_ = 0;
    free_.clear();
    live_.clear();
    if (size) free_[0] = size;
  }
  int64_t alloc(size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, sz = it->second;
      free_.erase(it);
      if (sz > bytes) free_[off + bytes] = sz - bytes;
      live_[off] = bytes;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return (int64_t)off;
    }
    return -1;
  }
  void release(int64_t off) {
    auto it = live_.find((size_t)off);
    if (it == live_.end()) return;
    size_t start = it->first, len = it->second;
    used_ -= len;
    live_.erase(it);
    auto nx = free_.lower_bound(start);
    if (nx != free_.end() && nx->first == start + len) {
      len += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == start) {
        start = pv->first;
        len += pv->second;
        free_.erase(pv);
      }
    }
    free_[start] = len;
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  std::map<size_t, size_t> live_;
};

struct Request {
  int64_t id = 0;
  int kind = FAP_REQ_GRAPH;
  int state = FAP_REQ_PENDING;
  int code = FAP_OK;
  std::string error;
  // graph
  int32_t V = 0;
  int64_t E = 0;
  std::vector<int32_t> row_offset;
  std::vector<int2> edges;
  int64_t block = -1;  // arena offset of the n x ld result
  int64_t ld = 0;
  // queries
  std::vector<int32_t> dst, uniq, qrow, answers;
  int64_t rows = -1;  // arena offset of uniq.size() x V distance rows
  // bookkeeping
  int64_t wave = -1;
  int32_t wave_size = 0;
  int readers = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;

  size_t stage_need(bool indexed) const {
    if (kind == FAP_REQ_GRAPH)
      return align_up(4 * ((size_t)V + 1)) + align_up(8 * (size_t)E) + sizeof(fapk::DenseJob) +
             sizeof(fapk::BatchJob) + 3 * sizeof(fapk::BatchSeg) * (size_t)((V + 63) / 64 + 1);
    const size_t q = dst.size(), u = indexed ? 0 : uniq.size();
    return 4 * u + 8 * u + 12 * q;
  }
  size_t arena_need(int32_t resident_V, bool indexed) const {
    if (kind == FAP_REQ_GRAPH) return 4 * (size_t)V * (size_t)ld;
    return indexed ? 0 : 4 * uniq.size() * (size_t)resident_V;
  }
};

// Slack for the alignment of a wave's descriptor arrays in staging.
constexpr size_t kStageSlack = 8 * kAlign;

// Query index over the resident graph (fap_serve_attach_graph with
// query_parts > 0).
struct QueryIndex {
  bool ready = false;
  int4 *pinfo = nullptr;   // K+1: {offset, size, boundary count, first B row}
  int64_t *moff = nullptr; // K+1: offset of each part matrix in M
  int32_t *M = nullptr;    // closed part matrices
  int32_t *B = nullptr;    // boundary SSSP rows (sum of boundary counts x V)
  int64_t bytes = 0;
  double build_ms = 0;
  void release() {
    void *bufs[] = {pinfo, moff, M, B};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
    pinfo = nullptr;
    moff = nullptr;
    M = nullptr;
    B = nullptr;
    bytes = 0;
    ready = false;
  }
};

struct Lane {
  fap_ctx *ctx = nullptr;
  char *h_stage = nullptr;  // pinned host
  char *d_stage = nullptr;  // device
  hipEvent_t done = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;  // big graphs beside the batched closure
  bool busy = false;
  int kind = FAP_REQ_GRAPH;
  std::vector<Request *> wave;
  size_t ans_off = 0;
};

}  // namespace
}  // namespace fapi

struct fap_server {
  int device = 0;
  fap_serve_opts opts{};
  std::vector<fapi::Lane> lanes;
  char *arena_base = nullptr;
  fapi::Arena arena;
  fap_graph_dev *resident = nullptr;
  fapi::QueryIndex index;
  // client-side device work (rows, checksums) runs on its own stream
  hipStream_t client = nullptr;
  unsigned long long *d_acc = nullptr;
  std::mutex client_mu;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::list<fapi::Request *> pending;
  std::unordered_map<int64_t, fapi::Request *> reqs;
  int64_t next_id = 1, next_wave = 0;
  uint64_t kick = 0;  // bumped by submit / release / stop
  bool stop = false;
  std::string fatal;
  // test hook (environment FAP_SERVE_FAIL_MIXED_WAVE=n): the n-th mixed
  // wave fails on the host after its big graphs were queued on the
  // auxiliary stream, before the join -- the failure path the lane must
  // survive without handing back blocks still being written
  int fail_mixed_wave = 0;
  std::thread thr;
  fapi::Clock::time_point t0;
  fap_serve_stats st{};
  double now_us() const {
    return std::chrono::duration<double, std::micro>(fapi::Clock::now() - t0).count();
  }
};

namespace fapi {
namespace {

void finish(fap_server *S, Request *r, int state, int code, const std::string &err) {
  r->state = state;
  r->code = code;
  r->error = err;
  r->t_done = S->now_us();
  if (state == FAP_REQ_FAILED) {
    S->st.failed++;
    if (r->block >= 0) S->arena.release(r->block);
    r->block = -1;
  } else {
    S->st.completed++;
  }
  if (r->rows >= 0) S->arena.release(r->rows);
  r->rows = -1;
  // the inputs were staged; free them
  std::vector<int32_t>().swap(r->row_offset);
  std::vector<int2>().swap(r->edges);
}

// Picks the next wave for lane L (lock held).  Returns true if anything
// changed (requests admitted or failed).
bool select_wave(fap_server *S, Lane &L) {
  if (S->pending.empty()) return false;
  const int kind = S->pending.front()->kind;
  const size_t stage_cap = (size_t)S->opts.stage_bytes - kStageSlack;
  const int32_t rV = S->resident ? S->resident->V : 0;
  size_t stage = 0;
  int64_t sources = 0;
  bool changed = false;
  L.wave.clear();
  for (auto it = S->pending.begin(); it != S->pending.end();) {
    if ((int)L.wave.size() >= S->opts.max_batch) break;
    Request *r = *it;
    if (r->kind != kind) {
      ++it;
      continue;
    }
    if (kind == FAP_REQ_QUERY && !S->resident) {
      finish(S, r, FAP_REQ_FAILED, FAP_E_STATE, "no resident graph attached");
      it = S->pending.erase(it);
      changed = true;
      continue;
    }
    const bool indexed = S->index.ready;
    const size_t sn = r->stage_need(indexed);
    const int64_t un = kind == FAP_REQ_QUERY && !indexed ? (int64_t)r->uniq.size() : 0;
    if (sn > stage_cap || un > S->opts.max_sources) {
      finish(S, r, FAP_REQ_FAILED, FAP_E_NOMEM,
             "request larger than one wave (staging bytes or sources per wave)");
      it = S->pending.erase(it);
      changed = true;
      continue;
    }
    if (stage + sn > stage_cap || sources + un > S->opts.max_sources) break;
    const size_t an = r->arena_need(rV, indexed);
    if (an > S->arena.size()) {
      finish(S, r, FAP_REQ_FAILED, FAP_E_NOMEM, "request larger than the HBM budget");
      it = S->pending.erase(it);
      changed = true;
      continue;
    }
    if (an > 0) {
      const int64_t off = S->arena.alloc(an);
      if (off < 0) break;  // wait for memory; later requests of this kind wait too
      if (kind == FAP_REQ_GRAPH) r->block = off;
      else r->rows = off;
    }
    stage += sn;
    sources += un;
    r->state = FAP_REQ_RUNNING;
    r->t_admit = S->now_us();
    L.wave.push_back(r);
    it = S->pending.erase(it);
    changed = true;
  }
  L.kind = kind;
  if (!L.wave.empty()) {
    const int64_t w = S->next_wave++;
    for (Request *r : L.wave) {
      r->wave = w;
      r->wave_size = (int32_t)L.wave.size();
    }
  }
  return changed;
}

// Stages and launches lane L's wave (lock not held; the wave's requests are
// RUNNING, so no client call touches them).
int launch_wave(fap_server *S, Lane &L) {
  char *h = L.h_stage;
  size_t off = 0;
  auto put = [&](const void *src, size_t bytes) {
    const size_t o = off;
    if (bytes) memcpy(h + o, src, bytes);
    off = align_up(o + bytes);
    return o;
  };
  hipStream_t s = L.ctx->stream;
  FAP_HIP(hipSetDevice(S->device));
  if (L.kind == FAP_REQ_GRAPH) {
    const int32_t cnt = (int32_t)L.wave.size();
    std::vector<size_t> o_ro(cnt), o_e(cnt);
    std::vector<int32_t *> mats(cnt);
    std::vector<int32_t> ns(cnt);
    std::vector<int64_t> lds(cnt);
    for (int32_t j = 0; j < cnt; j++) {
      Request *r = L.wave[j];
      o_ro[j] = put(r->row_offset.data(), 4 * ((size_t)r->V + 1));
      o_e[j] = put(r->edges.data(), 8 * (size_t)r->E);
      mats[j] = (int32_t *)(S->arena_base + r->block);
      ns[j] = r->V;
      lds[j] = r->ld;
    }
    std::vector<fapk::DenseJob> dj;
    int64_t cells = 0, rows = 0;
    for (int32_t j = 0; j < cnt; j++) {
      dj.push_back(fapk::DenseJob{(const int32_t *)(L.d_stage + o_ro[j]),
                                  (const int2 *)(L.d_stage + o_e[j]), mats[j], lds[j], cells,
                                  (int32_t)rows, ns[j]});
      cells += (int64_t)ns[j] * ns[j];
      rows += ns[j];
    }
    // Graphs of at least DENSE_PATH_N vertices are closed one by one with
    // the dense blocked FW (one such closure fills the GPU, and its fused
    // diagonal and single-kind launches close n = 4096 in 3.5 ms against
    // 6.1 ms in the batched closure); smaller ones share one batched
    // closure, where they run side by side.
    std::vector<int32_t *> bmats;
    std::vector<int32_t> bns;
    std::vector<int64_t> blds;
    std::vector<int32_t> big;
    for (int32_t j = 0; j < cnt; j++) {
      if (ns[j] >= DENSE_PATH_N) {
        big.push_back(j);
      } else {
        bmats.push_back(mats[j]);
        bns.push_back(ns[j]);
        blds.push_back(lds[j]);
      }
    }
    FwBatchDesc fd;
    build_fw_batch((int32_t)bmats.size(), bmats.data(), bns.data(), blds.data(), &fd);
    const size_t o_dj = put(dj.data(), sizeof(fapk::DenseJob) * dj.size());
    const size_t o_bj = put(fd.jobs.data(), sizeof(fapk::BatchJob) * fd.jobs.size());
    const size_t o_sg = put(fd.segs.data(), sizeof(fapk::BatchSeg) * fd.segs.size());
    if (off > (size_t)S->opts.stage_bytes) return fail(FAP_E_NOMEM, "wave staging overflow");
    FAP_HIP(hipMemcpyAsync(L.d_stage, h, off, hipMemcpyHostToDevice, s));
    FAP_TRY(launch_dense_jobs(L.ctx, (const fapk::DenseJob *)(L.d_stage + o_dj), cnt, cells, rows,
                              s));
    // With both kinds in the wave, the big graphs run on the lane
    // context's auxiliary stream beside the batched closure of the small
    // ones; the wave's completion (recorded on s) waits for both.
    const bool split = !big.empty() && !bmats.empty();
    hipStream_t sb = split ? L.ctx->aux_stream : s;
    if (split) {
      FAP_HIP(hipEventRecord(L.fork, s));
      FAP_HIP(hipStreamWaitEvent(sb, L.fork, 0));
    }
    for (int32_t j : big) FAP_TRY(launch_fw(L.ctx, mats[j], ns[j], lds[j], sb));
    if (split && S->fail_mixed_wave > 0 && --S->fail_mixed_wave == 0)
      return fail(FAP_E_STATE, "injected failure of a mixed wave after the fork");
    if (!bmats.empty())
      FAP_TRY(launch_fw_batch(L.ctx, fd, (const fapk::BatchJob *)(L.d_stage + o_bj),
                              (const fapk::BatchSeg *)(L.d_stage + o_sg), s));
    if (split) {
      FAP_HIP(hipEventRecord(L.join, sb));
      FAP_HIP(hipStreamWaitEvent(s, L.join, 0));
    }
  } else if (S->index.ready) {
    std::vector<int32_t> qs, qd;
    for (Request *r : L.wave)
      for (size_t q = 0; q < r->dst.size(); q++) {
        qs.push_back(r->uniq[r->qrow[q]]);
        qd.push_back(r->dst[q]);
      }
    const size_t o_qs = put(qs.data(), 4 * qs.size());
    const size_t o_qd = put(qd.data(), 4 * qd.size());
    const size_t staged = off;
    L.ans_off = off;
    off = align_up(off + 4 * qd.size());
    if (off > (size_t)S->opts.stage_bytes) return fail(FAP_E_NOMEM, "wave staging overflow");
    FAP_HIP(hipMemcpyAsync(L.d_stage, h, staged, hipMemcpyHostToDevice, s));
    const int nq = (int)qd.size();
    const fap_graph_dev *g = S->resident;
    hipLaunchKernelGGL(fapk::indexed_answers_kernel, dim3((nq + 63) / 64), dim3(64), 0, s,
                       g->graph_id, g->ed2st, S->index.pinfo, S->index.moff, S->index.M,
                       S->index.B, (int64_t)g->V, (const int32_t *)(L.d_stage + o_qs),
                       (const int32_t *)(L.d_stage + o_qd), (int32_t *)(L.d_stage + L.ans_off),
                       nq);
    FAP_CHECK_LAUNCH();
    FAP_HIP(hipMemcpyAsync(h + L.ans_off, L.d_stage + L.ans_off, 4 * (size_t)nq,
                           hipMemcpyDeviceToHost, s));
  } else {
    const int32_t V = S->resident->V;
    std::vector<int32_t> src, qrow, qdst;
    std::vector<int32_t *> rowptr;
    for (Request *r : L.wave) {
      const int32_t base = (int32_t)src.size();
      int32_t *rows = (int32_t *)(S->arena_base + r->rows);
      for (size_t k = 0; k < r->uniq.size(); k++) {
        src.push_back(r->uniq[k]);
        rowptr.push_back(rows + (int64_t)k * V);
      }
      for (size_t q = 0; q < r->dst.size(); q++) {
        qrow.push_back(base + r->qrow[q]);
        qdst.push_back(r->dst[q]);
      }
    }
    const size_t o_src = put(src.data(), 4 * src.size());
    const size_t o_rp = put(rowptr.data(), 8 * rowptr.size());
    const size_t o_qr = put(qrow.data(), 4 * qrow.size());
    const size_t o_qd = put(qdst.data(), 4 * qdst.size());
    const size_t staged = off;
    L.ans_off = off;
    off = align_up(off + 4 * qdst.size());
    if (off > (size_t)S->opts.stage_bytes) return fail(FAP_E_NOMEM, "wave staging overflow");
    FAP_HIP(hipMemcpyAsync(L.d_stage, h, staged, hipMemcpyHostToDevice, s));
    FAP_TRY(run_sssp_rows(L.ctx, S->resident, (const int32_t *)(L.d_stage + o_src),
                          (int32_t)src.size(), nullptr, 0, (int32_t *const *)(L.d_stage + o_rp),
                          s));
    const int nq = (int)qdst.size();
    hipLaunchKernelGGL(fapk::answers_kernel, dim3((nq + 255) / 256), dim3(256), 0, s,
                       (int32_t *const *)(L.d_stage + o_rp), (const int32_t *)(L.d_stage + o_qr),
                       (const int32_t *)(L.d_stage + o_qd), (int32_t *)(L.d_stage + L.ans_off),
                       nq);
    FAP_CHECK_LAUNCH();
    FAP_HIP(hipMemcpyAsync(h + L.ans_off, L.d_stage + L.ans_off, 4 * (size_t)nq,
                           hipMemcpyDeviceToHost, s));
  }
  FAP_HIP(hipEventRecord(L.done, s));
  return FAP_OK;
}

void complete_wave(fap_server *S, Lane &L) {
  const int32_t *ans = (const int32_t *)(L.h_stage + L.ans_off);
  for (Request *r : L.wave) {
    if (r->kind == FAP_REQ_QUERY) {
      r->answers.assign(ans, ans + r->dst.size());
      ans += r->dst.size();
      if (!S->index.ready) S->st.sssp_sources += (int64_t)r->uniq.size();
    } else {
      S->st.fw_ops += 2 * (int64_t)r->V * r->V * r->V;
    }
    finish(S, r, FAP_REQ_DONE, FAP_OK, "");
  }
  L.wave.clear();
  L.busy = false;
}

void fail_wave(fap_server *S, Lane &L, int code, const std::string &err) {
  for (Request *r : L.wave) finish(S, r, FAP_REQ_FAILED, code, err);
  L.wave.clear();
  L.busy = false;
}

void serve_loop(fap_server *S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  uint64_t seen = S->kick;
  for (;;) {
    bool progress = false, busy = false;
    for (Lane &L : S->lanes) {
      if (!L.busy) continue;
      const hipError_t e = hipEventQuery(L.done);
      if (e == hipErrorNotReady) {
        busy = true;
        continue;
      }
      if (e == hipSuccess) {
        complete_wave(S, L);
      } else {
        (void)hipGetLastError();
        S->fatal = std::string("device error in a wave: ") + hipGetErrorString(e);
        fail_wave(S, L, FAP_E_HIP, S->fatal);
      }
      progress = true;
    }
    if (S->stop && !busy) break;
    if (!S->stop && S->fatal.empty()) {
      for (Lane &L : S->lanes) {
        if (L.busy || S->pending.empty()) continue;
        if (select_wave(S, L)) progress = true;
        if (L.wave.empty()) continue;
        L.busy = true;
        S->st.waves++;
        S->st.max_wave = std::max<int32_t>(S->st.max_wave, (int32_t)L.wave.size());
        lk.unlock();
        const int rc = launch_wave(S, L);
        const std::string err = rc == FAP_OK ? std::string() : std::string(fap_last_error());
        if (rc != FAP_OK) {
          // a failure after the fork of a mixed wave (or inside a closure's
          // look-ahead) can leave kernels queued on the auxiliary stream
          // with no join back to the lane's stream: both must be idle before
          // fail_wave hands the wave's arena blocks back
          (void)hipStreamSynchronize(L.ctx->stream);
          (void)hipStreamSynchronize(L.ctx->aux_stream);
          (void)hipGetLastError();
        }
        lk.lock();
        if (rc != FAP_OK) fail_wave(S, L, rc, err);
        busy = true;
      }
    }
    if (progress) {
      S->cv_done.notify_all();
      continue;
    }
    if (busy) {
      S->cv_work.wait_for(lk, std::chrono::microseconds(S->opts.poll_us),
                          [&] { return S->kick != seen; });
    } else {
      S->cv_work.wait(lk, [&] { return S->kick != seen; });
    }
    seen = S->kick;
  }
  for (Request *r : S->pending) finish(S, r, FAP_REQ_FAILED, FAP_E_STATE, "server stopped");
  S->pending.clear();
  S->cv_done.notify_all();
}

int check_csr(int32_t V, int64_t E, const int32_t *ro, const int32_t *col, const int32_t *w) {
  if (V <= 0 || E < 0 || E > INT32_MAX) return fail(FAP_E_ARG, "bad V/E");
  if (!ro || (E > 0 && (!col || !w))) return fail(FAP_E_ARG, "null CSR array");
  if (ro[0] != 0 || ro[V] != E) return fail(FAP_E_ARG, "row_offset must start at 0 and end at E");
  for (int32_t v = 0; v < V; v++)
    if (ro[v + 1] < ro[v]) return fail(FAP_E_ARG, "row_offset not monotone");
  for (int64_t e = 0; e < E; e++) {
    if (col[e] < 0 || col[e] >= V) return fail(FAP_E_ARG, "col out of range");
    if (w[e] < 0 || w[e] >= FAP_INF) return fail(FAP_E_ARG, "weight outside [0, INF)");
  }
  return FAP_OK;
}

int64_t enqueue(fap_server *S, Request *r) {
  std::lock_guard<std::mutex> lk(S->mu);
  r->id = S->next_id++;
  r->t_submit = S->now_us();
  S->reqs[r->id] = r;
  S->pending.push_back(r);
  S->st.submitted++;
  S->kick++;
  S->cv_work.notify_all();
  return r->id;
}

void fill_info(const Request *r, fap_serve_info *info) {
  if (!info) return;
  info->state = r->state;
  info->kind = r->kind;
  info->n = r->kind == FAP_REQ_GRAPH ? r->V : (int32_t)r->dst.size();
  info->wave_size = r->wave_size;
  info->wave = r->wave;
  info->ld = r->ld;
  info->submit_us = r->t_submit;
  info->admit_us = r->t_admit;
  info->done_us = r->t_done;
}

// Looks up a finished graph request and pins it against release.
int pin_graph(fap_server *S, int64_t id, Request **out) {
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FAP_E_ARG, "unknown request id");
  Request *r = it->second;
  if (r->kind != FAP_REQ_GRAPH) return fail(FAP_E_ARG, "not a graph request");
  if (r->state != FAP_REQ_DONE) return fail(FAP_E_BUSY, "request not finished");
  r->readers++;
  *out = r;
  return FAP_OK;
}

void unpin(fap_server *S, Request *r) {
  std::lock_guard<std::mutex> lk(S->mu);
  r->readers--;
}

// The query index over the resident graph g (uploaded, same CSR as `csr`):
// partition into K parts, preCondition, SSSP rows from every boundary
// vertex, every part's matrix gathered and closed (one batched closure).
int build_index(fap_server *S, const HostCSR &csr, int32_t K) {
  fap_ctx *ctx = S->lanes[0].ctx;
  fap_graph_dev *g = S->resident;
  hipStream_t s = ctx->stream;
  QueryIndex &ix = S->index;
  const auto t0 = Clock::now();
  std::vector<int32_t> gid;
  bisection_partition(csr, K, &gid);
  PartInfo pi;
  precondition(csr, K, gid, &pi);
  FAP_TRY(fap_graph_set_partition(ctx, g, gid.data(), pi.st2ed.data(), pi.ed2st.data()));
  const int64_t V = csr.V;
  std::vector<int4> pinfo(K + 1, make_int4(0, 0, 0, 0));
  std::vector<int64_t> moff(K + 1, 0);
  std::vector<int32_t> srcs;
  int64_t nb = 0, mwords = 0;
  for (int32_t p = 1; p <= K; p++) {
    const int32_t off = pi.ver_off[p], n = pi.ver_num[p], b = pi.bdy_num[p];
    pinfo[p] = make_int4(off, n, b, (int32_t)nb);
    moff[p] = mwords;
    mwords += (int64_t)n * (((int64_t)n + 3) & ~(int64_t)3);
    for (int32_t k = 0; k < b; k++) srcs.push_back(pi.st2ed[off + k]);
    nb += b;
  }
  if (nb > INT32_MAX) return fail(FAP_E_ARG, "query index: too many boundary vertices");
  const size_t b_bytes = (size_t)std::max<int64_t>(nb, 1) * V * 4;
  const size_t m_bytes = (size_t)std::max<int64_t>(mwords, 1) * 4;
  hipError_t e = hipMalloc((void **)&ix.B, b_bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&ix.M, m_bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&ix.pinfo, sizeof(int4) * (K + 1));
  if (e == hipSuccess) e = hipMalloc((void **)&ix.moff, sizeof(int64_t) * (K + 1));
  int32_t *d_src = nullptr;
  if (e == hipSuccess) e = hipMalloc((void **)&d_src, sizeof(int32_t) * std::max<int64_t>(nb, 1));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (d_src) (void)hipFree(d_src);
    ix.release();
    return fail(FAP_E_NOMEM, std::string("query index allocation: ") + hipGetErrorString(e));
  }
  ix.bytes = (int64_t)(b_bytes + m_bytes);
  int rc = FAP_OK;
  e = hipMemcpyAsync(ix.pinfo, pinfo.data(), sizeof(int4) * (K + 1), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ix.moff, moff.data(), sizeof(int64_t) * (K + 1), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && nb)
    e = hipMemcpyAsync(d_src, srcs.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) rc = hip_fail(e, "query index upload", __FILE__, __LINE__);
  // stage 1: every boundary vertex's row
  if (rc == FAP_OK && nb) rc = run_sssp_rows(ctx, g, d_src, (int32_t)nb, ix.B, V, nullptr, s);
  // stage 2: part matrices, closed together
  std::vector<int32_t *> mats(K);
  std::vector<int32_t> ns(K);
  std::vector<int64_t> lds(K);
  for (int32_t p = 1; p <= K && rc == FAP_OK; p++) {
    const int4 q = pinfo[p];
    mats[p - 1] = ix.M + moff[p];
    ns[p - 1] = q.y;
    lds[p - 1] = ((int64_t)q.y + 3) & ~(int64_t)3;
    rc = launch_gather(g, q.x, q.y, q.z, ix.B + (int64_t)q.w * V, V, mats[p - 1], lds[p - 1], s);
  }
  fap_fw_plan *plan = nullptr;
  if (rc == FAP_OK) rc = fw_plan_create(ctx, K, mats.data(), ns.data(), lds.data(), 0, &plan);
  if (rc == FAP_OK) rc = fw_plan_run(plan, s);
  e = hipStreamSynchronize(s);
  if (plan) fw_plan_destroy(plan);
  (void)hipFree(d_src);
  if (rc == FAP_OK && e != hipSuccess) rc = hip_fail(e, "query index build", __FILE__, __LINE__);
  if (rc != FAP_OK) {
    ix.release();
    return rc;
  }
  ix.build_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  ix.ready = true;
  return FAP_OK;
}

void free_server(fap_server *S) {
  // the resident graph and its index first: freeing the graph uses lane 0's
  // context, destroyed below
  for (Lane &L : S->lanes)
    if (L.ctx) (void)hipStreamSynchronize(L.ctx->stream);
  S->index.release();
  if (S->resident) (void)fap_graph_free(S->lanes.empty() ? nullptr : S->lanes[0].ctx, S->resident);
  S->resident = nullptr;
  for (Lane &L : S->lanes) {
    if (L.ctx) (void)hipStreamSynchronize(L.ctx->stream);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.fork) (void)hipEventDestroy(L.fork);
    if (L.join) (void)hipEventDestroy(L.join);
    if (L.h_stage) (void)hipHostFree(L.h_stage);
    if (L.d_stage) (void)hipFree(L.d_stage);
    if (L.ctx) (void)fap_ctx_destroy(L.ctx);
  }
  if (S->client) {
    (void)hipStreamSynchronize(S->client);
    (void)hipStreamDestroy(S->client);
  }
  if (S->d_acc) (void)hipFree(S->d_acc);
  if (S->arena_base) (void)hipFree(S->arena_base);
  for (auto &kv : S->reqs) delete kv.second;
  delete S;
}

}  // namespace
}  // namespace fapi

using namespace fapi;

extern "C" {

void fap_serve_default_opts(fap_serve_opts *o) {
  if (!o) return;
  o->hbm_budget = (int64_t)8 << 30;
  o->stage_bytes = (int64_t)64 << 20;
  o->lanes = 3;
  o->max_batch = 64;
  o->max_sources = 256;
  o->poll_us = 20;
  o->query_parts = 0;
}

int fap_serve_create(int device, const fap_serve_opts *opts, fap_server **out) {
  if (!out) return fail(FAP_E_ARG, "null output");
  fap_serve_opts o;
  fap_serve_default_opts(&o);
  if (opts) o = *opts;
  if (o.hbm_budget < (1 << 20) || o.stage_bytes < (1 << 20) || o.lanes < 1 || o.lanes > 8 ||
      o.max_batch < 1 || o.max_sources < 1 || o.poll_us < 1 || o.query_parts < 0)
    return fail(FAP_E_ARG, "bad server options");
  o.stage_bytes = (int64_t)align_up((size_t)o.stage_bytes);
  fap_server *S = new fap_server();
  S->device = device;
  S->opts = o;
  S->t0 = Clock::now();
  if (const char *f = getenv("FAP_SERVE_FAIL_MIXED_WAVE")) S->fail_mixed_wave = atoi(f);
  S->lanes.resize(o.lanes);
  int rc = FAP_OK;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc((void **)&S->arena_base, (size_t)o.hbm_budget);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->client, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&S->d_acc, 64);
  for (int i = 0; i < o.lanes && e == hipSuccess && rc == FAP_OK; i++) {
    Lane &L = S->lanes[i];
    rc = fap_ctx_create(device, &L.ctx);
    if (rc != FAP_OK) break;
    e = hipHostMalloc((void **)&L.h_stage, (size_t)o.stage_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&L.d_stage, (size_t)o.stage_bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.join, hipEventDisableTiming);
  }
  if (e != hipSuccess || rc != FAP_OK) {
    if (e != hipSuccess) {
      (void)hipGetLastError();
      rc = fail(FAP_E_NOMEM, std::string("server allocation failed: ") + hipGetErrorString(e));
    }
    free_server(S);
    return rc;
  }
  S->arena.init((size_t)o.hbm_budget);
  S->st.arena_bytes = o.hbm_budget;
  S->st.lanes = o.lanes;
  S->thr = std::thread(serve_loop, S);
  *out = S;
  return FAP_OK;
}

int fap_serve_destroy(fap_server *S) {
  if (!S) return FAP_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->kick++;
  }
  S->cv_work.notify_all();
  if (S->thr.joinable()) S->thr.join();
  free_server(S);
  return FAP_OK;
}

int fap_serve_attach_graph(fap_server *S, int32_t V, int64_t E, const int32_t *ro,
                           const int32_t *col, const int32_t *w) {
  if (!S) return fail(FAP_E_ARG, "null server");
  FAP_TRY(check_csr(V, E, ro, col, w));
  std::lock_guard<std::mutex> lk(S->mu);
  for (auto &kv : S->reqs)
    if (kv.second->kind == FAP_REQ_QUERY &&
        (kv.second->state == FAP_REQ_PENDING || kv.second->state == FAP_REQ_RUNNING))
      return fail(FAP_E_BUSY, "query requests in flight");
  fap_graph_dev *g = nullptr;
  FAP_TRY(fap_graph_upload(S->lanes[0].ctx, V, E, ro, col, w, &g));
  // size every lane's SSSP state for any wave now, so that no workspace
  // grows (and synchronizes the device) while waves run
  size_t ws_bytes = 0;
  for (int32_t n = 1; n <= S->opts.max_sources; n++)
    ws_bytes = std::max(ws_bytes, sssp_ws_bytes(S->lanes[0].ctx, g, n));
  for (Lane &L : S->lanes) {
    const int rc = ensure_ws(L.ctx, L.ctx->ws_queue, ws_bytes);
    if (rc != FAP_OK) {
      (void)fap_graph_free(S->lanes[0].ctx, g);
      return rc;
    }
  }
  if (S->resident) {
    S->index.release();
    (void)fap_graph_free(S->lanes[0].ctx, S->resident);
  }
  S->resident = g;
  if (S->opts.query_parts > 0) {
    HostCSR csr;
    csr.V = V;
    csr.E = E;
    csr.row_offset.assign(ro, ro + V + 1);
    csr.col.assign(col, col + E);
    csr.w.assign(w, w + E);
    const int32_t K = std::min<int32_t>(S->opts.query_parts, V);
    const int rc = build_index(S, csr, K);
    if (rc != FAP_OK) {
      (void)fap_graph_free(S->lanes[0].ctx, S->resident);
      S->resident = nullptr;
      return rc;
    }
  }
  return FAP_OK;
}

int fap_serve_submit_graph(fap_server *S, int32_t V, int64_t E, const int32_t *ro,
                           const int32_t *col, const int32_t *w, int64_t *id) {
  if (!S || !id) return fail(FAP_E_ARG, "null argument");
  FAP_TRY(check_csr(V, E, ro, col, w));
  if ((int64_t)V * (((int64_t)V + 3) & ~3) * 4 > S->opts.hbm_budget)
    return fail(FAP_E_NOMEM, "graph result larger than the HBM budget");
  Request *r = new Request();
  r->kind = FAP_REQ_GRAPH;
  r->V = V;
  r->E = E;
  r->ld = ((int64_t)V + 3) & ~(int64_t)3;
  r->row_offset.assign(ro, ro + V + 1);
  r->edges.resize((size_t)E);
  for (int64_t e = 0; e < E; e++) r->edges[e] = make_int2(col[e], w[e]);
  *id = enqueue(S, r);
  return FAP_OK;
}

int fap_serve_submit_queries(fap_server *S, int32_t count, const int32_t *src, const int32_t *dst,
                             int64_t *id) {
  if (!S || !id || !src || !dst) return fail(FAP_E_ARG, "null argument");
  if (count < 1) return fail(FAP_E_ARG, "empty query batch");
  int32_t V;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->resident) return fail(FAP_E_STATE, "no resident graph attached");
    V = S->resident->V;
  }
  Request *r = new Request();
  r->kind = FAP_REQ_QUERY;
  r->dst.assign(dst, dst + count);
  std::unordered_map<int32_t, int32_t> slot;
  r->qrow.resize(count);
  for (int32_t q = 0; q < count; q++) {
    if (src[q] < 0 || src[q] >= V || dst[q] < 0 || dst[q] >= V) {
      delete r;
      return fail(FAP_E_ARG, "query vertex out of range at pair " + std::to_string(q));
    }
    auto it = slot.find(src[q]);
    if (it == slot.end()) {
      it = slot.emplace(src[q], (int32_t)r->uniq.size()).first;
      r->uniq.push_back(src[q]);
    }
    r->qrow[q] = it->second;
  }
  *id = enqueue(S, r);
  return FAP_OK;
}

int fap_serve_wait(fap_server *S, int64_t id, int32_t timeout_ms, fap_serve_info *info) {
  if (!S) return fail(FAP_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FAP_E_ARG, "unknown request id");
  Request *r = it->second;
  if (timeout_ms > 0)
    S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                        [&] { return r->state >= FAP_REQ_DONE; });
  fill_info(r, info);
  if (r->state == FAP_REQ_FAILED) return fail(r->code != FAP_OK ? r->code : FAP_E_STATE, r->error);
  if (r->state != FAP_REQ_DONE) return fail(FAP_E_BUSY, "request not finished");
  return FAP_OK;
}

int fap_serve_answers(fap_server *S, int64_t id, int32_t *out) {
  if (!S || !out) return fail(FAP_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FAP_E_ARG, "unknown request id");
  Request *r = it->second;
  if (r->kind != FAP_REQ_QUERY) return fail(FAP_E_ARG, "not a query request");
  if (r->state != FAP_REQ_DONE) return fail(FAP_E_BUSY, "request not finished");
  memcpy(out, r->answers.data(), 4 * r->answers.size());
  return FAP_OK;
}

int fap_serve_rows(fap_server *S, int64_t id, int32_t row0, int32_t nrows, int32_t *out) {
  if (!S || !out) return fail(FAP_E_ARG, "null argument");
  Request *r = nullptr;
  FAP_TRY(pin_graph(S, id, &r));
  int rc = FAP_OK;
  if (row0 < 0 || nrows < 0 || (int64_t)row0 + nrows > r->V) {
    rc = fail(FAP_E_ARG, "rows outside the result");
  } else if (nrows > 0) {
    std::lock_guard<std::mutex> ck(S->client_mu);
    const int32_t *src = (const int32_t *)(S->arena_base + r->block) + (int64_t)row0 * r->ld;
    hipError_t e = hipSetDevice(S->device);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(out, 4 * (size_t)r->V, src, 4 * (size_t)r->ld, 4 * (size_t)r->V,
                           (size_t)nrows, hipMemcpyDeviceToHost, S->client);
    if (e == hipSuccess) e = hipStreamSynchronize(S->client);
    if (e != hipSuccess) rc = hip_fail(e, "serve rows", __FILE__, __LINE__);
  }
  unpin(S, r);
  return rc;
}

int fap_serve_checksum(fap_server *S, int64_t id, uint64_t *sum, uint64_t *weighted) {
  if (!S || !sum || !weighted) return fail(FAP_E_ARG, "null argument");
  Request *r = nullptr;
  FAP_TRY(pin_graph(S, id, &r));
  int rc = FAP_OK;
  {
    std::lock_guard<std::mutex> ck(S->client_mu);
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipSetDevice(S->device);
    if (e == hipSuccess) rc = launch_fill_i32((int32_t *)S->d_acc, 4, 0, S->client);
    if (e == hipSuccess && rc == FAP_OK) {
      const int64_t cells = (int64_t)r->V * r->V;
      const int blocks = (int)std::min<int64_t>((cells + 255) / 256, 2048);
      hipLaunchKernelGGL(fapk::checksum2d_kernel, dim3(blocks), dim3(256), 0, S->client,
                         (const uint32_t *)(S->arena_base + r->block), r->ld, r->V, S->d_acc);
      e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(h, S->d_acc, sizeof(h), hipMemcpyDeviceToHost, S->client);
      if (e == hipSuccess) e = hipStreamSynchronize(S->client);
    }
    if (e != hipSuccess) rc = hip_fail(e, "serve checksum", __FILE__, __LINE__);
    *sum = h[0];
    *weighted = h[1];
  }
  unpin(S, r);
  return rc;
}

int fap_serve_block(fap_server *S, int64_t id, int32_t **d_block, int64_t *ld, int32_t *n) {
  if (!S || !d_block || !ld || !n) return fail(FAP_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FAP_E_ARG, "unknown request id");
  Request *r = it->second;
  if (r->kind != FAP_REQ_GRAPH) return fail(FAP_E_ARG, "not a graph request");
  if (r->state != FAP_REQ_DONE) return fail(FAP_E_BUSY, "request not finished");
  *d_block = (int32_t *)(S->arena_base + r->block);
  *ld = r->ld;
  *n = r->V;
  return FAP_OK;
}

int fap_serve_release(fap_server *S, int64_t id) {
  if (!S) return fail(FAP_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FAP_E_ARG, "unknown request id");
  Request *r = it->second;
  if (r->state == FAP_REQ_PENDING || r->state == FAP_REQ_RUNNING)
    return fail(FAP_E_BUSY, "request still queued or running");
  if (r->readers > 0) return fail(FAP_E_BUSY, "request being read");
  if (r->block >= 0) S->arena.release(r->block);
  if (r->rows >= 0) S->arena.release(r->rows);
  S->reqs.erase(it);
  delete r;
  S->kick++;
  S->cv_work.notify_all();
  return FAP_OK;
}

int fap_serve_get_stats(fap_server *S, fap_serve_stats *out) {
  if (!S || !out) return fail(FAP_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  out->pending = (int64_t)S->pending.size();
  int64_t running = 0;
  for (const Lane &L : S->lanes) running += (int64_t)L.wave.size();
  out->running = running;
  out->arena_used = (int64_t)S->arena.used();
  out->arena_peak = (int64_t)S->arena.peak();
  out->index_bytes = S->index.bytes;
  out->index_build_ms = S->index.build_ms;
  if (!S->fatal.empty()) return fail(FAP_E_HIP, S->fatal);
  return FAP_OK;
}

}  // extern "C"
