// batch.hip -- many uneven closures, phase-synchronous launches.
//
// The reference solves parts one after another (fap/fap.cpp:189-193) and
// spreads them over processes with MPI (fap/fap_more_task_mpi.cpp:34-91).
// Here a batch of matrices (the K parts of one graph, or independent graphs)
// is closed together:
//
//  * each matrix (job) is cut into 64 x 64 tiles and its blocked
//    Floyd-Warshall becomes a sequence of phases, each a set of independent
//    tile tasks.  Pivot rounds go in pairs: diag(r), panels(r), cross(r)
//    (row and column r+1 with pivot r), diag(r+1), panels(r+1), rest2(r)
//    (every tile off rows/columns r, r+1, both pivots in one 128-deep
//    product: C is read and written once per two rounds), cross(r) (row and
//    column r with pivot r+1, after rest2 has read them); an unpaired last
//    round is diag, panels, rest.  Min-plus is exact, so the order of the
//    two pivots' updates changes no bit;
//  * phase p of every job goes into launch p: a queue of tasks, largest
//    jobs first, consumed by a persistent grid (workgroup w takes tasks w,
//    w + grid, ...).  The queue is described by one segment per job phase
//    (its first task and count), found by binary search, so descriptors
//    grow with the number of phases, not tiles.  Tasks of one launch never
//    depend on each other; the kernel boundary orders a job's phase p
//    after its phase p-1;
//  * the launch sequence of a batch is a plan: descriptors uploaded once,
//    launches captured in a hipGraph and replayed.
//
// The phase launches have no wait inside a kernel: every workgroup runs
// its tasks and exits.  fw_queue_kernel runs the same tasks in ONE launch,
// with bounded in-kernel waits on per-job completion counters; its
// workgroups claim tasks in queue order, which is what lets every wait end
// (an earlier single-launch variant dealt tasks statically, so resident
// workgroups could wait on tasks of workgroups not yet dispatched, and it
// made no progress on the device).
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "batch_desc.h"
#include "common.h"
#include "minplus_tile.h"

namespace fapk {

// Launches of diagonal tasks only: one 1024-thread workgroup per task, a
// thread per 1 x 4 strip (lds_fw64, as the dense diagonal kernel), about
// twice as fast per block as the 256-thread closure.
__global__ __launch_bounds__(1024) void fw_diag_batch_kernel(const BatchJob *__restrict__ jobs,
                                                            const BatchSeg *__restrict__ segs,
                                                            int nseg, int ntasks) {
  __shared__ __attribute__((aligned(16))) int32_t D[BLK_LDS_WORDS];
  const int t = threadIdx.x;
  for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segs[mid].first <= ti) lo = mid;
      else hi = mid - 1;
    }
    const BatchSeg sg = segs[lo];
    const BatchJob job = jobs[sg.job];
    const int base = (sg.kind_r >> 3) * BT;
    const int kb = imin(BT, job.n - base);
    int32_t *blk = job.mat + (int64_t)base * job.ld + base;
    for (int idx = t; idx < 64 * 64; idx += 1024) {
      const int i = idx >> 6, j = idx & 63;
      D[i * BLK_LD + j] = (i < kb && j < kb) ? blk[(int64_t)i * job.ld + j] : kInf;
    }
    __syncthreads();
    lds_fw64<BLK_LD>(D, 0, kb);
    for (int idx = t; idx < 64 * 64; idx += 1024) {
      const int i = idx >> 6, j = idx & 63;
      if (i < kb && j < kb) blk[(int64_t)i * job.ld + j] = D[i * BLK_LD + j];
    }
    __syncthreads();  // the LDS tile is reused by the next task
  }
}

// The segment holding task ti: segments tile [0, ntasks) in order.
template <typename Seg>
__device__ __forceinline__ int find_seg(const Seg *__restrict__ segs, int nseg, int ti) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].first <= ti) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Task idx of a panel, single-pivot (rest and cross kinds) or pair-pass
// phase with pivot block r of `job`: one 64 x 64 tile update.
__device__ __forceinline__ void batch_tile_task(const BatchJob &job, int kind, int r, int idx,
                                                int32_t *lds) {
  const int nb = job.nb, n = job.n;
  const int64_t ld = job.ld;
  int32_t *d = job.mat;
  const int base = r * BT;
  const int kb = imin(BT, n - base);
  const int flags =
      ((((uintptr_t)d) & 15) == 0 && (ld & 3) == 0) ? (MP_VEC_A | MP_VEC_B | MP_VEC_C) : 0;
  if (kind == TASK_PANEL) {
    const int32_t *diag = d + (int64_t)base * ld + base;
    int I, J;
    panel_tile(r, nb, idx, &I, &J);
    if (I == r) {
      int32_t *c = d + (int64_t)base * ld + (int64_t)J * BT;
      MpSmall::run(diag, ld, c, ld, c, ld, kb, imin(BT, n - J * BT), kb, true, flags, lds);
    } else {
      int32_t *c = d + (int64_t)I * BT * ld + base;
      MpSmall::run(c, ld, diag, ld, c, ld, imin(BT, n - I * BT), kb, kb, true, flags, lds);
    }
  } else if (kind == TASK_REST2) {
    // pivots r and r + 1 at once: blocks r, r + 1 are adjacent, so the two
    // A tiles of row I form one 64 x 128 panel and the two B tiles of
    // column J one 128 x 64 panel
    int I, J;
    pair_tile(r, nb, idx, &I, &J);
    MpSmall::run(d + (int64_t)I * BT * ld + base, ld, d + (int64_t)base * ld + (int64_t)J * BT,
                 ld, d + (int64_t)I * BT * ld + (int64_t)J * BT, ld, imin(BT, n - I * BT),
                 imin(BT, n - J * BT), imin(2 * BT, n - base), true, flags, lds);
  } else {
    // one tile: off row/column r with pivot r, or of row/column q with
    // pivot p (cross kinds)
    int I, J, p;
    single_tile(kind, r, nb, idx, &I, &J, &p);
    const int pb = p * BT;
    MpSmall::run(d + (int64_t)I * BT * ld + pb, ld, d + (int64_t)pb * ld + (int64_t)J * BT,
                 ld, d + (int64_t)I * BT * ld + (int64_t)J * BT, ld, imin(BT, n - I * BT),
                 imin(BT, n - J * BT), imin(BT, n - pb), true, flags, lds);
  }
}

// Panel, single-pivot (rest and cross kinds) and pair-pass tasks; every
// task of a launch runs on one instance (KIND = kind_instance of its kind:
// the phase patterns keep launches homogeneous and the host checks), so
// each instance fits 64 VGPRs: 8 waves per SIMD, which hides the memory
// waits of these short (64- or 128-deep) products better than larger
// per-lane state would (measured: resident A panels at 4 waves, and four
// 8 x 8-lane tiles sharing A at 4 waves, were no faster although they read
// fewer bytes).
template <int KIND>
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(8))) void fw_phase_kernel(
    const BatchJob *__restrict__ jobs, const BatchSeg *__restrict__ segs, int nseg, int ntasks) {
  static_assert(KIND == TASK_PANEL || KIND == TASK_REST || KIND == TASK_REST2, "kernel instance");
  __shared__ __attribute__((aligned(16))) int32_t lds[MpSmall::LDS_WORDS];
  for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
    const BatchSeg sg = segs[find_seg(segs, nseg, ti)];
    const int kind = KIND == TASK_REST ? (sg.kind_r & 7) : KIND;
    batch_tile_task(jobs[sg.job], kind, sg.kind_r >> 3, ti - sg.first, lds);
    __syncthreads();  // the LDS buffer is reused by the next task
  }
}

// XCD (accelerator complex die) the calling wave runs on: the XCC_ID
// hardware register.  Each XCD has its own L2.
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15; }

// Bits of the XCDs a grid's workgroups run on.
__global__ void xcc_probe_kernel(int32_t *mask) {
  if (threadIdx.x == 0) atomicOr(mask, 1 << xcc_id());
}

// Every phase of every job in ONE launch.  The jobs are split over the
// device's XCDs, one queue per XCD: a queue is the phase-ordered launch
// sequence of its jobs laid end to end, and a workgroup serves the queue
// of the XCD it runs on (qmap[xcc_id()]), claiming tasks in queue order
// from that queue's counter, each claim made while the workgroup runs (the
// next one while its current task computes).  A task first waits until
// `need` tasks of its job have completed (done[job]): all of them come
// earlier in the same queue, so each was claimed by a workgroup that is
// running and waits only for tasks earlier still -- the first unfinished
// task of a queue can always run, whatever the grid and however many
// workgroups are resident.  The wait is bounded: past `limit` wall-clock
// ticks, or once another workgroup has given up (ctrl[0] != 0), the
// workgroup raises ctrl[0] and exits, and so does every other, so the
// launch always ends; the host reports the flag, and checks every job's
// count (a queue whose XCD ran no workgroup would be left unfinished).
// Every task of a job runs on one XCD, so writer and reader share that
// XCD's L2: a task's stores are complete in L2 before its count is raised
// (s_waitcnt vmcnt(0) in every wave, then the barrier), with no L2
// write-back; the next task's thread 0 takes an agent-scope acquire after
// its wait (dropping stale lines of the CU's L1; an L1-only invalidate,
// buffer_inv sc0, left stale lines and wrong results), and the barrier
// orders every wave's loads after it.  Measured on the 512-graph batch:
// agent-scope release and acquire in every wave 158 ms, XCD-local release
// with every wave acquiring 38 ms, thread 0 acquiring 23 ms.  Diagonal
// tasks close their block with 256 threads (close_block64).  All task
// kinds in one kernel take 80 VGPRs (64 spill): 6 waves per SIMD.
//   qinfo: [0, 16) qmap (XCC id -> queue or -1); then per queue q, at
//          16 + 4q: first segment, segment count, task count, first record
//   all_rec: per task {job, kind_r, index in its phase, need}, filled by
//          fw_queue_map_kernel (one load per claimed task)
//   ctrl:  [0] flag; claim counter of queue q at 32 (q + 1) (own cache line)
__global__ __launch_bounds__(MP_THREADS) __attribute__((amdgpu_waves_per_eu(6))) void fw_queue_kernel(
    const BatchJob *__restrict__ jobs, const int32_t *__restrict__ qinfo,
    const int4 *__restrict__ all_rec, int32_t *ctrl, int32_t *done, int64_t limit) {
  constexpr int W = BLK_LDS_WORDS > MpSmall::LDS_WORDS ? BLK_LDS_WORDS : MpSmall::LDS_WORDS;
  __shared__ __attribute__((aligned(16))) int32_t lds[W];
  __shared__ int32_t s_next, s_go;
  const int t = threadIdx.x;
  const int q = qinfo[xcc_id()];
  if (q < 0) return;
  const int ntasks = qinfo[16 + 4 * q + 2];
  const int4 *rec = all_rec + qinfo[16 + 4 * q + 3];
  int32_t *claim = ctrl + 32 * (q + 1);
  if (t == 0) s_next = atomicAdd(claim, 1);
  __syncthreads();
  int ti = s_next;
  int nxt = ntasks;  // thread 0: the next claim, made while the current task runs
  while (ti < ntasks) {
    const int4 tr = rec[ti];  // {job, kind_r, index, need}
    if (t == 0) {
      int go = 1;
      // the flag is read only while waiting: a task run after another
      // workgroup gave up is harmless, the call reports the flag anyway
      if (tr.w > 0) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(&done[tr.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tr.w) {
          if (__hip_atomic_load(&ctrl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
              (int64_t)(__builtin_amdgcn_s_memrealtime() - t0) > limit) {
            go = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      if (!go) {
        __hip_atomic_store(&ctrl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        // published after the task, so the claim's round trip overlaps it
        nxt = atomicAdd(claim, 1);
        // the CU's L1 may hold lines of this job's tiles from before their
        // last writes: drop them (an agent-scope acquire; the barrier below
        // orders every wave's loads after it)
        if (tr.w > 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      s_go = go;
    }
    __syncthreads();
    if (!s_go) break;  // thread 0 writes s_go again only after the barrier below
    const BatchJob job = jobs[tr.x];
    const int kind = tr.y & 7, r = tr.y >> 3;
    if (kind == TASK_DIAG) {
      const int base = r * BT;
      close_block64(job.mat + (int64_t)base * job.ld + base, job.ld, imin(BT, job.n - base), lds);
    } else {
      batch_tile_task(job, kind, r, tr.z, lds);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete in L2
    if (t == 0) s_next = nxt;
    __syncthreads();                // ... every wave's; the LDS buffer is free again
    if (t == 0) __hip_atomic_fetch_add(&done[tr.x], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ti = s_next;  // thread 0 writes it again only after the next task
  }
}

// Task records of every queue: workgroup b writes segment b's (segments of
// all queues, each queue's records from its own offset).
__global__ void fw_queue_map_kernel(const BatchQSeg *__restrict__ all_segs,
                                    const int32_t *__restrict__ seg_rec_off, int nseg,
                                    int4 *rec) {
  for (int b = blockIdx.x; b < nseg; b += gridDim.x) {
    const BatchQSeg sg = all_segs[b];
    int4 *m = rec + seg_rec_off[b] + sg.first;
    for (int i = threadIdx.x; i < sg.count; i += blockDim.x)
      m[i] = make_int4(sg.job, sg.kind_r, i, sg.need);
  }
}

// Clears a queue launch's flag, claim counters and per-job counters (a
// kernel, not a memset, so the sequence can be captured).
__global__ void fw_queue_reset_kernel(int32_t *ctrl, int nctrl, int32_t *done, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nctrl) ctrl[i] = 0;
  if (i < count) done[i] = 0;
}

}  // namespace fapk

namespace fapi {

// Launch p of a batch holds phase p of every job that has one, jobs in
// decreasing size (longest processing time first), each job's phase one
// segment of consecutive tasks.  A one-block job has only its diagonal.
void build_fw_batch(int32_t count, int32_t *const *mats, const int32_t *n, const int64_t *ld,
                    FwBatchDesc *out) {
  out->jobs.resize(count);
  out->segs.clear();
  out->seg_off.assign(1, 0);
  out->launch_tasks.clear();
  out->ops = 0;
  std::vector<int32_t> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
  // Job j's phase p goes to launch p.  (Offsetting some jobs' phases, so a
  // launch mixes one job's diagonal with another's remaining tiles,
  // measured 3% slower.)  Pair and tail patterns align kind for kind
  // (diag, panels, then cross or rest), so every launch runs one instance.
  int launches = 0;
  std::vector<std::vector<int32_t>> ph(count);  // kind | (r << 3) per phase
  for (int i = 0; i < count; i++) {
    const int j = order[i];
    const int nb = (n[j] + fapk::BT - 1) / fapk::BT;
    out->jobs[j] = fapk::BatchJob{mats[j], ld[j], n[j], nb};
    out->ops += 2 * (int64_t)n[j] * n[j] * n[j];
    if (n[j] <= 0) continue;
    std::vector<int32_t> &pj = ph[j];
    auto add = [&](int kind, int r) { pj.push_back(kind | (r << 3)); };
    if (nb == 1) add(fapk::TASK_DIAG, 0);
    for (int r = 0; nb > 1 && r < nb;) {
      add(fapk::TASK_DIAG, r);
      add(fapk::TASK_PANEL, r);
      if (r + 1 < nb) {
        add(fapk::TASK_CROSS_A, r);
        add(fapk::TASK_DIAG, r + 1);
        add(fapk::TASK_PANEL, r + 1);
        add(fapk::TASK_REST2, r);
        add(fapk::TASK_CROSS_B, r);
        r += 2;
      } else {
        add(fapk::TASK_REST, r);
        r += 1;
      }
    }
    launches = std::max(launches, (int)pj.size());
  }
  for (int L = 0; L < launches; L++) {
    int32_t first = 0;
    for (int j : order) {
      if (L >= (int)ph[j].size()) continue;
      const int kind_r = ph[j][L];
      const int cnt = fapk::batch_task_count(kind_r & 7, out->jobs[j].nb);
      if (cnt <= 0) continue;
      out->segs.push_back(fapk::BatchSeg{j, kind_r, first, cnt});
      first += cnt;
    }
    out->launch_tasks.push_back(first);
    out->seg_off.push_back((int32_t)out->segs.size());
  }
  out->key = fap_tune_key_batch(count, n);
}

int32_t build_fw_queue(const FwBatchDesc &desc, std::vector<fapk::BatchQSeg> *out) {
  out->clear();
  std::vector<int32_t> before(desc.jobs.size(), 0);  // tasks of each job queued so far
  int32_t first = 0;
  for (size_t L = 0; L < desc.launch_tasks.size(); L++) {
    // a launch holds at most one phase per job: its tasks need all of the
    // job's tasks queued before this launch
    for (int q = desc.seg_off[L]; q < desc.seg_off[L + 1]; q++) {
      const fapk::BatchSeg &sg = desc.segs[q];
      out->push_back(fapk::BatchQSeg{sg.job, sg.kind_r, first + sg.first, sg.count,
                                     before[sg.job], 0});
    }
    for (int q = desc.seg_off[L]; q < desc.seg_off[L + 1]; q++)
      before[desc.segs[q].job] += desc.segs[q].count;
    first += desc.launch_tasks[L];
  }
  return first;
}

// XCDs of the context's device (bit mask), probed once: a grid of a few
// workgroups per CU records the XCC ids it runs on.
static int device_xcc_mask(fap_ctx *ctx, hipStream_t s, int *mask) {
  if (ctx->xcc_mask <= 0) {
    int32_t *d = nullptr, h = 0;
    FAP_HIP(hipMalloc(&d, 256));
    hipLaunchKernelGGL(fapk::fw_queue_reset_kernel, dim3(1), dim3(64), 0, s, d, 1,
                       (int32_t *)nullptr, 0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
      hipLaunchKernelGGL(fapk::xcc_probe_kernel, dim3(std::max(ctx->num_cus, 1) * 4), dim3(64), 0,
                         s, d);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "XCD probe", __FILE__, __LINE__);
    if (h <= 0) return fail(FAP_E_HIP, "fw_batched_persistent: XCD probe saw no XCC id");
    ctx->xcc_mask = h;
  }
  *mask = ctx->xcc_mask;
  return FAP_OK;
}

// Host side of a persistent launch (fw_queue_kernel): jobs split over the
// device's XCDs by work (2 n^3, largest first, each to the least loaded),
// one queue per XCD; one device block holds
//   jobs | segments | qinfo | each segment's first task record (uploaded)
//   | ctrl | done | task records.
int build_queue_block(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                      const int64_t *ld, hipStream_t s, QueueBlock *qb) {
  // tasks of one phase write distinct tiles, and a tile's rows must not
  // share a 128-byte cache line with another tile's
  for (int32_t j = 0; j < count; j++)
    if (n[j] > 0 && ((((uintptr_t)d_mats[j]) & 127) != 0 || (ld[j] & 31) != 0))
      return fail(FAP_E_ARG, "fw_batched_persistent: matrix " + std::to_string(j) +
                                 " needs 128-byte aligned rows (base address a multiple of 128, "
                                 "ld a multiple of 32)");
  FAP_HIP(hipSetDevice(ctx->device));
  int khz = 0;
  FAP_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (khz <= 0) return fail(FAP_E_HIP, "fw_batched_persistent: no wall clock rate");
  qb->wait_us = ctx->opt_batch_wait_us > 0 ? ctx->opt_batch_wait_us : 2000000;
  qb->limit = qb->wait_us * khz / 1000;
  int mask = 0;
  FAP_TRY(device_xcc_mask(ctx, s, &mask));
  std::vector<int> xccs;
  for (int x = 0; x < 16; x++)
    if (mask & (1 << x)) xccs.push_back(x);
  const int nq = (int)std::min<size_t>(xccs.size(), (size_t)count);

  std::vector<int32_t> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  std::vector<double> load(nq, 0.0);
  std::vector<std::vector<int32_t>> qjobs(nq);
  for (int j : order) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    qjobs[q].push_back(j);
    load[q] += 2.0 * n[j] * (double)n[j] * n[j] + 1.0;
  }
  // per queue: its own launch sequence laid end to end; jobs and segments
  // concatenated over the queues (segment job indices made global)
  std::vector<fapk::BatchJob> jobs;
  std::vector<fapk::BatchQSeg> segs;
  std::vector<int32_t> qinfo(16 + 4 * nq, 0), seg_rec_off;
  qb->job_total.clear();
  qb->ops = 0;
  for (int x = 0; x < 16; x++) qinfo[x] = -1;
  int64_t ntasks = 0;
  for (int q = 0; q < nq; q++) {
    qinfo[xccs[q]] = q;
    std::vector<int32_t *> m;
    std::vector<int32_t> nn;
    std::vector<int64_t> ll;
    for (int j : qjobs[q]) {
      m.push_back(d_mats[j]);
      nn.push_back(n[j]);
      ll.push_back(ld[j]);
    }
    FwBatchDesc d;
    build_fw_batch((int32_t)m.size(), m.data(), nn.data(), ll.data(), &d);
    qb->ops += d.ops;
    std::vector<fapk::BatchQSeg> qs;
    const int32_t qt = build_fw_queue(d, &qs);
    if (ntasks + qt + (int64_t)ctx->num_cus * 8 >= INT32_MAX)
      return fail(FAP_E_ARG, "fw_batched_persistent: too many tasks");
    const int32_t job_base = (int32_t)jobs.size();
    qinfo[16 + 4 * q] = (int32_t)segs.size();
    qinfo[16 + 4 * q + 1] = (int32_t)qs.size();
    qinfo[16 + 4 * q + 2] = qt;
    qinfo[16 + 4 * q + 3] = (int32_t)ntasks;
    for (const fapk::BatchJob &jb : d.jobs) {
      jobs.push_back(jb);
      qb->job_total.push_back(0);
    }
    for (fapk::BatchQSeg g : qs) {
      g.job += job_base;
      qb->job_total[g.job] += g.count;
      seg_rec_off.push_back((int32_t)ntasks);
      segs.push_back(g);
    }
    ntasks += qt;
  }
  qb->ntasks = ntasks;
  qb->njobs = (int)jobs.size();
  qb->nseg = (int)segs.size();
  qb->nctrl = 32 * (nq + 1);
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t jobs_b = sizeof(fapk::BatchJob) * jobs.size();
  qb->seg_off = up(jobs_b);
  qb->qi_off = up(qb->seg_off + sizeof(fapk::BatchQSeg) * segs.size());
  qb->sro_off = up(qb->qi_off + sizeof(int32_t) * qinfo.size());
  qb->ctrl_off = up(qb->sro_off + sizeof(int32_t) * segs.size());
  qb->done_off = qb->ctrl_off + sizeof(int32_t) * qb->nctrl;
  qb->map_off = up(qb->done_off + sizeof(int32_t) * std::max(qb->njobs, 1));
  qb->total = qb->map_off + sizeof(int4) * (size_t)std::max<int64_t>(ntasks, 1);
  qb->host.assign(qb->ctrl_off, 0);
  char *h = qb->host.data();
  if (jobs_b) memcpy(h, jobs.data(), jobs_b);
  if (!segs.empty()) {
    memcpy(h + qb->seg_off, segs.data(), sizeof(fapk::BatchQSeg) * segs.size());
    memcpy(h + qb->sro_off, seg_rec_off.data(), sizeof(int32_t) * segs.size());
  }
  memcpy(h + qb->qi_off, qinfo.data(), sizeof(int32_t) * qinfo.size());
  qb->grid = ctx->num_cus * 6;  // 6 resident workgroups per CU; dealt to the XCDs in turn
  return FAP_OK;
}

// Descriptors up and the task records filled (once per block).
int upload_queue_block(const QueueBlock &qb, char *dev, hipStream_t s) {
  FAP_HIP(hipMemcpyAsync(dev, qb.host.data(), qb.ctrl_off, hipMemcpyHostToDevice, s));
  if (qb.nseg > 0) {
    hipLaunchKernelGGL(fapk::fw_queue_map_kernel, dim3(std::min(qb.nseg, 65535)), dim3(256), 0, s,
                       (const fapk::BatchQSeg *)(dev + qb.seg_off),
                       (const int32_t *)(dev + qb.sro_off), qb.nseg, (int4 *)(dev + qb.map_off));
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

// One closure of the batch: counters cleared, then the persistent launch
// (kernels only, so the pair can be captured in a graph).
int enqueue_queue_run(const QueueBlock &qb, char *dev, hipStream_t s) {
  if (qb.ntasks <= 0) return FAP_OK;
  int32_t *ctrl = (int32_t *)(dev + qb.ctrl_off), *done = (int32_t *)(dev + qb.done_off);
  hipLaunchKernelGGL(fapk::fw_queue_reset_kernel, dim3((std::max(qb.nctrl, qb.njobs) + 255) / 256),
                     dim3(256), 0, s, ctrl, qb.nctrl, done, qb.njobs);
  FAP_CHECK_LAUNCH();
  hipLaunchKernelGGL(fapk::fw_queue_kernel, dim3(qb.grid), dim3(fapk::MP_THREADS), 0, s,
                     (const fapk::BatchJob *)dev, (const int32_t *)(dev + qb.qi_off),
                     (const int4 *)(dev + qb.map_off), ctrl, done, qb.limit);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// After a closure has finished: the give-up flag and every job's count.
int check_queue_run(const QueueBlock &qb, const char *dev) {
  if (qb.ntasks <= 0) return FAP_OK;
  std::vector<int32_t> got(qb.njobs + 1, 0);
  FAP_HIP(hipMemcpy(got.data(), dev + qb.done_off, sizeof(int32_t) * qb.njobs,
                    hipMemcpyDeviceToHost));
  FAP_HIP(hipMemcpy(got.data() + qb.njobs, dev + qb.ctrl_off, sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  if (got[qb.njobs] != 0)
    return fail(FAP_E_DEVICE, "fw_batched_persistent: a task waited longer than " +
                                  std::to_string((long long)qb.wait_us) +
                                  " us for its job's earlier phases; the launch gave up "
                                  "(matrices partly closed)");
  for (int j = 0; j < qb.njobs; j++)
    if (got[j] != qb.job_total[j])
      return fail(FAP_E_DEVICE, "fw_batched_persistent: a matrix's queue ran " +
                                    std::to_string(got[j]) + " of its " +
                                    std::to_string(qb.job_total[j]) +
                                    " tasks (no workgroup on its XCD)");
  return FAP_OK;
}

// The batch in one launch, synchronously.
static int run_fw_queue(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                        const int64_t *ld, hipStream_t s) {
  QueueBlock qb;
  FAP_TRY(build_queue_block(ctx, count, d_mats, n, ld, s, &qb));
  if (qb.ntasks <= 0) return FAP_OK;
  char *dev = nullptr;
  FAP_HIP(hipMalloc(&dev, qb.total));
  int rc = upload_queue_block(qb, dev, s);
  if (rc == FAP_OK) rc = enqueue_queue_run(qb, dev, s);
  if (rc == FAP_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "fw_batched_persistent", __FILE__, __LINE__);
  }
  if (rc == FAP_OK) rc = check_queue_run(qb, dev);
  (void)hipFree(dev);
  return rc;
}

int launch_fw_batch(fap_ctx *ctx, const FwBatchDesc &desc, const fapk::BatchJob *d_jobs,
                    const fapk::BatchSeg *d_segs, hipStream_t s, size_t first, size_t last) {
  const int cap = ctx->num_cus * batch_grid_per_cu(ctx, desc.key);
  last = std::min(last, desc.launch_tasks.size());
  for (size_t p = first; p < last; p++) {
    const int cnt = desc.launch_tasks[p];
    const int nseg = desc.seg_off[p + 1] - desc.seg_off[p];
    if (cnt <= 0 || nseg <= 0) continue;
    // one kernel instance for every task of the launch -> that instance
    int kind = fapk::kind_instance(desc.segs[desc.seg_off[p]].kind_r & 7);
    for (int q = desc.seg_off[p] + 1; q < desc.seg_off[p + 1]; q++)
      if (fapk::kind_instance(desc.segs[q].kind_r & 7) != kind) kind = -1;
    const dim3 grid(std::min(cnt, cap)), block(fapk::MP_THREADS);
    const fapk::BatchSeg *sg = d_segs + desc.seg_off[p];
    if (kind == fapk::TASK_DIAG)
      hipLaunchKernelGGL(fapk::fw_diag_batch_kernel, dim3(std::min(cnt, ctx->num_cus * 2)),
                         dim3(1024), 0, s, d_jobs, sg, nseg, cnt);
    else if (kind == fapk::TASK_PANEL)
      hipLaunchKernelGGL(fapk::fw_phase_kernel<fapk::TASK_PANEL>, grid, block, 0, s, d_jobs, sg,
                         nseg, cnt);
    else if (kind == fapk::TASK_REST)
      hipLaunchKernelGGL(fapk::fw_phase_kernel<fapk::TASK_REST>, grid, block, 0, s, d_jobs, sg,
                         nseg, cnt);
    else if (kind == fapk::TASK_REST2)
      hipLaunchKernelGGL(fapk::fw_phase_kernel<fapk::TASK_REST2>, grid, block, 0, s, d_jobs, sg,
                         nseg, cnt);
    else  // build_fw_batch never mixes instances in one launch
      return fail(FAP_E_STATE, "batched closure: launch " + std::to_string(p) + " mixes task kinds");
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

}  // namespace fapi

// A plan splits its jobs (size order, alternately) into up to two chains of
// launches on two streams: every job's phases stay in order on its chain,
// and the chains share no matrix.  The second chain starts half a round
// pair after the first, so one chain's short, latency-bound launches
// (diagonal blocks, panels, cross tiles) overlap the other's 128-deep
// passes instead of running alone.
constexpr int PLAN_CHAINS = 2;
constexpr int PLAN_CHAIN_MIN_JOBS = 8;  // fewer jobs: one chain
constexpr size_t PLAN_CHAIN_OFFSET = 3;  // diag, panels, cross of the first round

struct fap_fw_plan {
  fap_ctx *ctx = nullptr;
  int chains = 1;
  fapi::FwBatchDesc host[PLAN_CHAINS];
  void *desc[PLAN_CHAINS] = {};  // jobs | segments (device), per chain
  size_t segs_off[PLAN_CHAINS] = {};
  hipStream_t side[PLAN_CHAINS] = {};  // chains 1.. run on side[1..]
  hipEvent_t ev_join[PLAN_CHAINS] = {};
  hipEvent_t ev_fork = nullptr, ev_offset = nullptr;
  hipEvent_t ev_last = nullptr;  // recorded on the run's stream after each run
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t ops = 0;
  fapi::QueueBlock *queue = nullptr;  // persistent plan: one launch per run
  char *qdev = nullptr;               // its device block
  int64_t runs = 0;
};

namespace fapi {

static int chain_launch(fap_fw_plan *p, int c, hipStream_t s, size_t first = 0,
                        size_t last = (size_t)-1) {
  return launch_fw_batch(p->ctx, p->host[c], (const fapk::BatchJob *)p->desc[c],
                         (const fapk::BatchSeg *)((char *)p->desc[c] + p->segs_off[c]), s, first,
                         last);
}

static int plan_launch(fap_fw_plan *p, hipStream_t s) {
  if (p->queue) return enqueue_queue_run(*p->queue, p->qdev, s);
  if (p->chains == 1) return chain_launch(p, 0, s);
  // fork: the side stream starts after what precedes the plan on s, and
  // after the first chain's first PLAN_CHAIN_OFFSET launches; join: s waits
  // for the side chain
  FAP_HIP(hipEventRecord(p->ev_fork, s));
  for (int c = 1; c < p->chains; c++) FAP_HIP(hipStreamWaitEvent(p->side[c], p->ev_fork, 0));
  FAP_TRY(chain_launch(p, 0, s, 0, PLAN_CHAIN_OFFSET));
  FAP_HIP(hipEventRecord(p->ev_offset, s));
  FAP_TRY(chain_launch(p, 0, s, PLAN_CHAIN_OFFSET));
  for (int c = 1; c < p->chains; c++) {
    FAP_HIP(hipStreamWaitEvent(p->side[c], p->ev_offset, 0));
    FAP_TRY(chain_launch(p, c, p->side[c]));
    FAP_HIP(hipEventRecord(p->ev_join[c], p->side[c]));
    FAP_HIP(hipStreamWaitEvent(s, p->ev_join[c], 0));
  }
  return FAP_OK;
}

void fw_plan_destroy(fap_fw_plan *p) {
  if (!p) return;
  // the last run may still be in flight: wait for its end on the run's
  // stream (ev_last, which follows the join of every chain) and for the
  // plan's side streams before the exec and descriptors go; other streams
  // of the device are not waited for
  if (p->ev_last) (void)hipEventSynchronize(p->ev_last);
  for (int c = 0; c < PLAN_CHAINS; c++)
    if (p->side[c]) (void)hipStreamSynchronize(p->side[c]);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (int c = 0; c < PLAN_CHAINS; c++) {
    if (p->desc[c]) (void)hipFree(p->desc[c]);
    if (p->ev_join[c]) (void)hipEventDestroy(p->ev_join[c]);
    if (p->side[c]) (void)hipStreamDestroy(p->side[c]);
  }
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_offset) (void)hipEventDestroy(p->ev_offset);
  if (p->ev_last) (void)hipEventDestroy(p->ev_last);
  if (p->qdev) (void)hipFree(p->qdev);
  delete p->queue;
  delete p;
}

static int plan_upload(fap_fw_plan *p, int c) {
  const FwBatchDesc &h = p->host[c];
  const size_t jobs_b = sizeof(fapk::BatchJob) * h.jobs.size();
  p->segs_off[c] = (jobs_b + 255) & ~(size_t)255;
  const size_t segs_b = sizeof(fapk::BatchSeg) * h.segs.size();
  const size_t total = p->segs_off[c] + std::max<size_t>(segs_b, 16);
  std::vector<char> host(total, 0);
  if (jobs_b) memcpy(host.data(), h.jobs.data(), jobs_b);
  if (segs_b) memcpy(host.data() + p->segs_off[c], h.segs.data(), segs_b);
  hipError_t e = hipMalloc(&p->desc[c], total);
  if (e == hipSuccess) e = hipMemcpy(p->desc[c], host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "batch plan descriptors", __FILE__, __LINE__);
  return FAP_OK;
}

int fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *mats, const int32_t *n,
                   const int64_t *ld, int use_graph, fap_fw_plan **out) {
  fap_fw_plan *p = new fap_fw_plan();
  p->ctx = ctx;
  int rc = FAP_OK;
  hipError_t e = hipSuccess;
  p->chains = count >= PLAN_CHAIN_MIN_JOBS ? PLAN_CHAINS : 1;
  e = hipEventCreateWithFlags(&p->ev_last, hipEventDisableTiming);
  if (e != hipSuccess) {
    rc = hip_fail(e, "batch plan event", __FILE__, __LINE__);
    fw_plan_destroy(p);
    return rc;
  }
  if (p->chains == 1) {
    build_fw_batch(count, mats, n, ld, &p->host[0]);
  } else {
    std::vector<int32_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    std::vector<int32_t *> cm[PLAN_CHAINS];
    std::vector<int32_t> cn[PLAN_CHAINS];
    std::vector<int64_t> cl[PLAN_CHAINS];
    for (int i = 0; i < count; i++) {
      const int c = i % PLAN_CHAINS, j = order[i];
      cm[c].push_back(mats[j]);
      cn[c].push_back(n[j]);
      cl[c].push_back(ld[j]);
    }
    for (int c = 0; c < PLAN_CHAINS; c++)
      build_fw_batch((int32_t)cn[c].size(), cm[c].data(), cn[c].data(), cl[c].data(),
                     &p->host[c]);
    // both chains tune and report under the whole batch's bucket
    for (int c = 0; c < PLAN_CHAINS; c++) p->host[c].key = fap_tune_key_batch(count, n);
    e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_offset, hipEventDisableTiming);
    for (int c = 1; c < p->chains && e == hipSuccess; c++) {
      e = hipStreamCreateWithFlags(&p->side[c], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join[c], hipEventDisableTiming);
    }
    if (e != hipSuccess) rc = hip_fail(e, "batch plan streams", __FILE__, __LINE__);
  }
  for (int c = 0; c < p->chains && rc == FAP_OK; c++) {
    p->ops += p->host[c].ops;
    rc = plan_upload(p, c);
  }
  bool any = false;
  for (int c = 0; c < p->chains; c++) any = any || !p->host[c].segs.empty();
  if (rc == FAP_OK && use_graph && any) {
    hipStream_t cs = ctx->aux_stream;  // capture on a stream with no other work queued
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) rc = plan_launch(p, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &p->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess && rc == FAP_OK) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (rc == FAP_OK && e != hipSuccess) rc = hip_fail(e, "batch plan graph capture", __FILE__, __LINE__);
  }
  if (rc != FAP_OK) {
    fw_plan_destroy(p);
    return rc;
  }
  *out = p;
  return FAP_OK;
}

// A plan of one persistent launch per run: the queue block built and
// uploaded (task records filled) once; with use_graph the reset and
// the launch are captured.
int fw_plan_create_persistent(fap_ctx *ctx, int32_t count, int32_t *const *mats,
                              const int32_t *n, const int64_t *ld, int use_graph,
                              fap_fw_plan **out) {
  fap_fw_plan *p = new fap_fw_plan();
  p->ctx = ctx;
  p->queue = new QueueBlock();
  int rc = FAP_OK;
  hipError_t e = hipEventCreateWithFlags(&p->ev_last, hipEventDisableTiming);
  if (e != hipSuccess) rc = hip_fail(e, "batch plan event", __FILE__, __LINE__);
  if (rc == FAP_OK) rc = build_queue_block(ctx, count, mats, n, ld, ctx->stream, p->queue);
  if (rc == FAP_OK) {
    p->ops = p->queue->ops;
    e = hipMalloc(&p->qdev, p->queue->total);
    if (e != hipSuccess) rc = hip_fail(e, "persistent plan block", __FILE__, __LINE__);
  }
  if (rc == FAP_OK) rc = upload_queue_block(*p->queue, p->qdev, ctx->stream);
  if (rc == FAP_OK) {
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(e, "persistent plan upload", __FILE__, __LINE__);
  }
  if (rc == FAP_OK && use_graph && p->queue->ntasks > 0) {
    hipStream_t cs = ctx->aux_stream;  // capture on a stream with no other work queued
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) rc = plan_launch(p, cs);
    hipError_t e2 = hipStreamEndCapture(cs, &p->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess && rc == FAP_OK) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (rc == FAP_OK && e != hipSuccess) rc = hip_fail(e, "persistent plan graph capture", __FILE__, __LINE__);
  }
  if (rc != FAP_OK) {
    fw_plan_destroy(p);
    return rc;
  }
  *out = p;
  return FAP_OK;
}

int fw_plan_run(fap_fw_plan *p, hipStream_t s) {
  if (p->exec) {
    FAP_HIP(hipGraphLaunch(p->exec, s));
  } else {
    FAP_TRY(plan_launch(p, s));
  }
  FAP_HIP(hipEventRecord(p->ev_last, s));
  p->runs++;
  return FAP_OK;
}

// Scratch of fap_fw_batched_noalloc: the batch's job and segment
// descriptors, in a file-scope device array instead of a hipMalloc'd
// buffer.  Each device holds its own copy until the process ends.  In host
// code `noalloc_scratch` names a host-side stand-in, never the device
// memory: every host use goes through the address hipGetSymbolAddress
// returned for the current device (noalloc_scratch_ptr).
__device__ __align__(256) char noalloc_scratch[FAP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `noalloc_scratch` on `device` (made current by the
// caller), looked up once per device.
static int noalloc_scratch_ptr(int device, char **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(FAP_E_ARG, "no-alloc scratch: device id " + std::to_string(device) +
                               " beyond the " + std::to_string(kMaxDevices) + " supported");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    FAP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch)));
    if (!ptr) return fail(FAP_E_HIP, "no-alloc scratch: null symbol address");
    g_scratch_ptr[device] = ptr;
  }
  *out = (char *)g_scratch_ptr[device];
  return FAP_OK;
}

static int check_jobs(int32_t count, int32_t *const *d_mats, const int32_t *n, const int64_t *ld) {
  if (count < 0 || (count > 0 && (!d_mats || !n || !ld)))
    return fail(FAP_E_ARG, "fw_batched: bad arguments");
  int64_t tiles = 0;
  for (int32_t j = 0; j < count; j++) {
    if (n[j] < 0 || (n[j] > 0 && (!d_mats[j] || ld[j] < n[j])))
      return fail(FAP_E_ARG, "fw_batched: job " + std::to_string(j) + " has a bad size/stride");
    const int64_t nb = (n[j] + 63) / 64;
    tiles += nb * nb * nb;
  }
  if (tiles > ((int64_t)1 << 28)) return fail(FAP_E_ARG, "fw_batched: batch too large");
  return FAP_OK;
}

}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_fw_batched(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                   const int64_t *ld, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  FAP_TRY(check_jobs(count, d_mats, n, ld));
  if (count == 0) return FAP_OK;
  fap_fw_plan *p = nullptr;
  FAP_TRY(fw_plan_create(ctx, count, d_mats, n, ld, 0, &p));
  hipStream_t s = pick_stream(ctx, stream);
  int rc = fw_plan_run(p, s);
  if (rc == FAP_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "fw_batched", __FILE__, __LINE__);
  }
  fw_plan_destroy(p);
  return rc;
}

int fap_fw_batched_noalloc(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                           const int64_t *ld, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  FAP_TRY(check_jobs(count, d_mats, n, ld));
  if (count == 0) return FAP_OK;
  FAP_HIP(hipSetDevice(ctx->device));
  FwBatchDesc d;
  build_fw_batch(count, d_mats, n, ld, &d);
  const size_t jobs_b = sizeof(fapk::BatchJob) * d.jobs.size();
  const size_t segs_off = (jobs_b + 255) & ~(size_t)255;
  const size_t segs_b = sizeof(fapk::BatchSeg) * d.segs.size();
  if (segs_off + segs_b > (size_t)FAP_NOALLOC_SCRATCH_BYTES)
    return fail(FAP_E_ARG, "fw_batched_noalloc: the batch's descriptors need " +
                               std::to_string(segs_off + segs_b) + " bytes, above the " +
                               std::to_string((long long)FAP_NOALLOC_SCRATCH_BYTES) +
                               "-byte scratch (fewer or smaller matrices per call)");
  char *scratch = nullptr;
  FAP_TRY(noalloc_scratch_ptr(ctx->device, &scratch));
  hipStream_t s = pick_stream(ctx, stream);
  // the copies are ordered after earlier work on s, so an earlier call on
  // the same stream has finished reading the scratch before they land
  if (jobs_b) FAP_HIP(hipMemcpyAsync(scratch, d.jobs.data(), jobs_b, hipMemcpyHostToDevice, s));
  if (segs_b)
    FAP_HIP(hipMemcpyAsync(scratch + segs_off, d.segs.data(), segs_b, hipMemcpyHostToDevice, s));
  FAP_TRY(launch_fw_batch(ctx, d, (const fapk::BatchJob *)scratch,
                          (const fapk::BatchSeg *)(scratch + segs_off), s));
  FAP_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

int fap_fw_batched_persistent(fap_ctx *ctx, int32_t count, int32_t *const *d_mats,
                              const int32_t *n, const int64_t *ld, void *stream) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  FAP_TRY(check_jobs(count, d_mats, n, ld));
  if (count == 0) return FAP_OK;
  return run_fw_queue(ctx, count, d_mats, n, ld, pick_stream(ctx, stream));
}

int fap_fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                       const int64_t *ld, int use_graph, fap_fw_plan **out) {
  if (!ctx || !out) return fail(FAP_E_ARG, "null argument");
  FAP_TRY(check_jobs(count, d_mats, n, ld));
  return fw_plan_create(ctx, count, d_mats, n, ld, use_graph, out);
}

int fap_fw_plan_create_persistent(fap_ctx *ctx, int32_t count, int32_t *const *d_mats,
                                  const int32_t *n, const int64_t *ld, int use_graph,
                                  fap_fw_plan **out) {
  if (!ctx || !out) return fail(FAP_E_ARG, "null argument");
  FAP_TRY(check_jobs(count, d_mats, n, ld));
  return fw_plan_create_persistent(ctx, count, d_mats, n, ld, use_graph, out);
}

int fap_fw_plan_check(fap_fw_plan *plan) {
  if (!plan) return fail(FAP_E_ARG, "null plan");
  if (plan->runs == 0) return FAP_OK;
  FAP_HIP(hipEventSynchronize(plan->ev_last));
  if (plan->queue) return check_queue_run(*plan->queue, plan->qdev);
  return FAP_OK;
}

int fap_fw_plan_run(fap_fw_plan *plan, void *stream) {
  if (!plan) return fail(FAP_E_ARG, "null plan");
  return fw_plan_run(plan, pick_stream(plan->ctx, stream));
}

int fap_fw_plan_info(const fap_fw_plan *plan, int32_t *launches, int64_t *tasks, int64_t *ops) {
  if (!plan) return fail(FAP_E_ARG, "null plan");
  int64_t t = 0;
  int32_t l = 0;
  if (plan->queue) {
    t = plan->queue->ntasks;
    l = t > 0 ? 1 : 0;
  }
  for (int c = 0; c < plan->chains && !plan->queue; c++) {
    for (int32_t x : plan->host[c].launch_tasks) t += x;
    l += (int32_t)plan->host[c].launch_tasks.size();
  }
  if (launches) *launches = l;
  if (tasks) *tasks = t;
  if (ops) *ops = plan->ops;
  return FAP_OK;
}

int fap_fw_plan_destroy(fap_fw_plan *plan) {
  fw_plan_destroy(plan);
  return FAP_OK;
}

}  // extern "C"
