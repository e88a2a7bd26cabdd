// batch_desc.h -- descriptors shared by the batched closure (batch.hip),
// the dense matrix build (dense.hip) and the request loop (serve.hip).
//
// A batch is described by plain structs that the host writes once (into
// pinned or pageable memory) and copies to the device in one transfer; the
// launch functions take the device copies, so a caller that owns its own
// staging (the request loop) never waits on the host between launches.
#pragma once

#include <stdint.h>

#include <vector>

#include "common.h"

namespace fapk {

constexpr int BT = 64;  // batched closure tile

// One matrix of a batched closure.
struct BatchJob {
  int32_t *mat;
  int64_t ld;
  int32_t n;
  int32_t nb;  // ceil(n / BT)
};

// One phase of one job: tasks [first, first + count) of its launch are the
// tiles of that phase.  kind_r = kind | (pivot block << 3).
struct BatchSeg {
  int32_t job;
  int32_t kind_r;
  int32_t first;
  int32_t count;
};

// One phase of one job in the single-launch queue (fw_queue_kernel): tasks
// [first, first + count) of the whole queue, which start only once `need`
// tasks of the same job have completed -- every task of the job's earlier
// phases, all of which come earlier in the queue.
struct BatchQSeg {
  int32_t job;
  int32_t kind_r;
  int32_t first;
  int32_t count;
  int32_t need;
  int32_t pad;
};

// Phases of a job's blocked Floyd-Warshall.  Rounds go in pairs (r, r + 1):
//   DIAG(r), PANEL(r), CROSS_A(r), DIAG(r+1), PANEL(r+1), REST2(r), CROSS_B(r)
// where CROSS_A updates row and column r+1 with pivot r (what DIAG and PANEL
// of round r+1 read), REST2 updates every tile off rows/columns r, r+1 with
// both pivots in one 128-deep pass (C is read and written once for two
// rounds), and CROSS_B then updates row and column r with pivot r+1 (after
// REST2 has read them).  A last unpaired round is DIAG, PANEL, REST.  No
// task reads a tile that another task of the same phase writes.
enum : int {
  TASK_DIAG = 0,
  TASK_PANEL = 1,
  TASK_REST = 2,     // single pivot r, tiles off row/column r
  TASK_CROSS_A = 3,  // pivot r: row r+1 (J != r), column r+1 (I != r, r+1)
  TASK_CROSS_B = 4,  // pivot r+1: row r (J != r+1), column r (I != r, r+1)
  TASK_REST2 = 5,    // pivots r, r+1: tiles off rows/columns r, r+1
};

__host__ __device__ inline int batch_task_count(int kind, int nb) {
  switch (kind) {
    case TASK_DIAG: return 1;
    case TASK_PANEL: return 2 * (nb - 1);
    case TASK_REST: return (nb - 1) * (nb - 1);
    case TASK_CROSS_A:
    case TASK_CROSS_B: return 2 * nb - 3;
    default: return (nb - 2) * (nb - 2);
  }
}

// Tile u of a cross phase with pivot p on row/column q (|p - q| = 1): first
// row q without column p, then column q without rows p and q.
__host__ __device__ inline void cross_tile(int u, int nb, int p, int q, int *I, int *J) {
  if (u < nb - 1) {
    *I = q;
    *J = u < p ? u : u + 1;
  } else {
    const int v = u - (nb - 1), lo = p < q ? p : q;
    *I = v < lo ? v : v + 2;
    *J = q;
  }
}

// Task idx of a panel phase: row r without (r, r), then column r.
__host__ __device__ inline void panel_tile(int r, int nb, int idx, int *I, int *J) {
  if (idx < nb - 1) {
    *I = r;
    *J = idx < r ? idx : idx + 1;
  } else {
    const int u = idx - (nb - 1);
    *I = u < r ? u : u + 1;
    *J = r;
  }
}

// Task idx of a single-pivot phase: tile (I, J) and its pivot block p --
// rest: every tile off row/column r, pivot r; cross kinds: cross_tile.
__host__ __device__ inline void single_tile(int kind, int r, int nb, int idx, int *I, int *J,
                                            int *p) {
  if (kind == TASK_REST) {
    const int m = nb - 1, i = idx / m, j = idx - (idx / m) * m;
    *I = i + (i >= r);
    *J = j + (j >= r);
    *p = r;
  } else {
    *p = kind == TASK_CROSS_A ? r : r + 1;
    cross_tile(idx, nb, *p, kind == TASK_CROSS_A ? r + 1 : r, I, J);
  }
}

// Task idx of a pair pass: tile (I, J) off rows/columns r and r + 1.
__host__ __device__ inline void pair_tile(int r, int nb, int idx, int *I, int *J) {
  const int m = nb - 2, i = idx / m, j = idx - (idx / m) * m;
  *I = i < r ? i : i + 2;
  *J = j < r ? j : j + 2;
}

// Kernel instance that runs a task kind (the cross kinds run on the
// single-pivot instance); every task of a launch runs on one instance.
__host__ __device__ constexpr int kind_instance(int kind) {
  return kind == TASK_CROSS_A || kind == TASK_CROSS_B ? TASK_REST : kind;
}

// One graph's dense matrix: INF, arcs in CSR order, zero diagonal.
struct DenseJob {
  const int32_t *row_offset;
  const int2 *edges;
  int32_t *mat;
  int64_t ld;
  int64_t cell_base;  // prefix of n^2 over earlier jobs
  int32_t row_base;   // prefix of n over earlier jobs
  int32_t n;
};

}  // namespace fapk

namespace fapi {

// Host side of a batched closure: jobs, the segments of every launch (launch
// p = segments [seg_off[p], seg_off[p + 1])) and the task count per launch.
struct FwBatchDesc {
  std::vector<fapk::BatchJob> jobs;
  std::vector<fapk::BatchSeg> segs;
  std::vector<int32_t> seg_off;
  std::vector<int32_t> launch_tasks;
  int64_t ops = 0;
  int64_t key = 0;  // tune cache bucket (fap_tune_key_batch)
};

void build_fw_batch(int32_t count, int32_t *const *mats, const int32_t *n, const int64_t *ld,
                    FwBatchDesc *out);
// The launches of `desc` as one queue, launch after launch (so a task's
// own job's earlier phases always come before it); returns the task count.
int32_t build_fw_queue(const FwBatchDesc &desc, std::vector<fapk::BatchQSeg> *out);
// Host side of a batch's persistent launch (fw_queue_kernel): the
// descriptor bytes to upload (host, the first ctrl_off bytes of the device
// block) and the block's layout.
struct QueueBlock {
  std::vector<char> host;
  size_t seg_off = 0, qi_off = 0, sro_off = 0, ctrl_off = 0, done_off = 0, map_off = 0,
         total = 0;
  int nctrl = 0, njobs = 0, nseg = 0, grid = 0;
  int64_t ntasks = 0, ops = 0, limit = 0, wait_us = 0;
  std::vector<int32_t> job_total;  // tasks per job (block order)
};
int build_queue_block(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                      const int64_t *ld, hipStream_t s, QueueBlock *qb);
int upload_queue_block(const QueueBlock &qb, char *dev, hipStream_t s);
int enqueue_queue_run(const QueueBlock &qb, char *dev, hipStream_t s);
int check_queue_run(const QueueBlock &qb, const char *dev);

// Launch sequence of a batch whose jobs and segments are already on the
// device (same order as in `desc`).
int launch_fw_batch(fap_ctx *ctx, const FwBatchDesc &desc, const fapk::BatchJob *d_jobs,
                    const fapk::BatchSeg *d_segs, hipStream_t s,
                    size_t first = 0, size_t last = (size_t)-1);

// Multi-source SSSP (sssp.hip): source i's distances go to d_out + i * ld_out,
// or to d_rowptr[i] when d_rowptr is given (a device array of row pointers).
int run_sssp_rows(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *d_sources, int32_t n_src,
                  int32_t *d_out, int64_t ld_out, int32_t *const *d_rowptr, hipStream_t s);

// Device bytes of SSSP state run_sssp_rows needs for n_src sources.
size_t sssp_ws_bytes(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src);

// Part matrix gather (submat.hip) and batched-closure plans (batch.hip).
int launch_gather(const fap_graph_dev *g, int32_t start, int32_t n_sub, int32_t n_bdy,
                  const int32_t *rowblock, int64_t ld_rb, int32_t *sub, int64_t ld_sub,
                  hipStream_t s);
int fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *mats, const int32_t *n,
                   const int64_t *ld, int use_graph, fap_fw_plan **out);
int fw_plan_run(fap_fw_plan *p, hipStream_t s);
void fw_plan_destroy(fap_fw_plan *p);

// Dense build of jobs already on the device; `cells` = sum n^2, `rows` = sum n.
int launch_dense_jobs(fap_ctx *ctx, const fapk::DenseJob *d_jobs, int32_t njobs, int64_t cells,
                      int64_t rows, hipStream_t s);

}  // namespace fapi
