/*
 * fap_dev.h -- C ABI of the MI355X (gfx950) subgraph-solve path.
 *
 * Plain C: opaque handles, int32 data, explicit leading dimensions, device
 * pointers, a `void *stream` (hipStream_t; NULL = the context's own stream).
 * Every function returns FAP_OK (0) or a negative FAP_E* code; the message of
 * the last failure on the calling thread is fap_last_error().  Calls are
 * asynchronous on the stream unless stated otherwise.
 *
 * Numbers: distances and weights are int32 >= 0, INF = FAP_INF = INT32_MAX/2,
 * so INF + INF does not overflow.  The reference computes in float with
 * INF = 1e8f (fap/utils/parameter.h:26); results agree exactly while every
 * finite distance is < 2^24.
 *
 * Reference interfaces replaced (paths relative to the FastAPSP repository):
 *   fap_sssp_multi      <- handle_boundry_path      fap/kernel/batch_sssp.h:254-277
 *                          handle_boundry_AMD_GPU   fap/kernel/hip/hip_kernel.h:17-20
 *   fap_fw_blocked      <- floyd_path               fap/kernel/floyd.h:150-158
 *                          floyd_AMD_path           fap/kernel/hip/hip_kernel.h:24
 *   fap_minplus         <- min_plus_path_advanced   fap/kernel/minplus.h:78-89
 *                          minplus_AMD_path         fap/kernel/hip/hip_kernel.h:40-41
 *   fap_minplus_path    <- min_plus_path_CPU_advanced (with path) fap/kernel/minplus.h:51-65
 *   fap_fw_blocked_path <- floyd_CPU_path (with path)  fap/kernel/floyd.h:38-49
 *   fap_gather_submat   <- MysubMatBuild_path       fap/utils/util_Myalgorithm.h:30-73
 *   (A operand view)    <- MyMat1Build              fap/utils/util_Myalgorithm.h:88-95
 *   fap_scatter_submat  <- MysubMatDecode_path      fap/utils/util_Myalgorithm.h:75-86
 *   fap_subgraph_solve  <- fapGraph::run            fap/fap.cpp:92-181
 *   fap_batch_*         <- the MPI multi-task driver fap/fap_more_task_mpi.cpp:34-91,
 *                          collapsed into one device work queue on one GPU
 *   fap_graph_build_index  (stage 1 of batched solves) handle_boundry_path's rows,
 *                          fap/kernel/batch_sssp.h:254-277, by another exact route
 */
#ifndef FAP_DEV_H_
#define FAP_DEV_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FAP_INF 0x3FFFFFFF

enum {
  FAP_OK = 0,
  FAP_E_ARG = -1,     /* invalid argument (sizes, null pointers, ranges) */
  FAP_E_HIP = -2,     /* HIP runtime error (message has the HIP string) */
  FAP_E_NOMEM = -3,   /* device or host allocation failed / over budget */
  FAP_E_STATE = -4,   /* call out of order (e.g. solve before preCondition) */
  FAP_E_IO = -5,      /* file missing or malformed */
  FAP_E_DEVICE = -6,  /* a device-side bounded wait gave up (reported, not hung) */
  FAP_E_BUSY = -7,    /* not finished yet, or still in use (serving: wait / release) */
};

typedef struct fap_ctx fap_ctx;
typedef struct fap_graph_dev fap_graph_dev;

/* Last error message of the calling thread ("" if none). */
const char *fap_last_error(void);
/* Library version string. */
const char *fap_version(void);

/* ---- context, memory and stream helpers (synchronous unless noted) ---- */
int fap_device_count(int *count);
int fap_ctx_create(int device, fap_ctx **out);
int fap_ctx_destroy(fap_ctx *ctx);
int fap_ctx_stream(fap_ctx *ctx, void **stream);
/* Tuning options; they change speed, never results.  "Auto" (0, or -1 for
 * fw_lookahead; the defaults) defers to the context's tune cache
 * (fap_tune.h), then to the default in brackets:
 *   "fw_lookahead" -1/0/1   overlap FW round r+1's diagonal/panels with round r [0]
 *   "fw_tile"      0/64/128 FW block size [64]
 *   "sssp_threads" 0/256/512/1024  threads per SSSP workgroup [1024 while the sources
 *                           fit on the CUs, 512 up to twice that, 256 beyond]
 *   "sssp_delta"   >= 0     near-far band width [graph's own, else 8 / 6 / 4 x mean
 *                           weight for sources up to the CU count / twice it / more]
 *   "sssp_group"   0/1/4/8  sources per SSSP workgroup pass [1]
 *   "sssp_bin"     0..3     reorder SSSP near queues by vertex block when they hold at
 *                           least 2x / 1x / 0.5x the workgroup's threads (1 / 2 / 3);
 *                           0 off [3]
 *   "sssp_pack"    0/1/2    SSSP per-vertex state: distance and queue tag in one 64-bit
 *                           workspace word (1), or distance and two queue bits in one
 *                           32-bit word of the output row, with a third less
 *                           workspace (2) [1]
 *   "sssp_wgs"     0..8     SSSP workgroups per CU [1024 threads per CU]
 *   "sssp_arcs"    0/1/2/4  arcs each lane relaxes per SSSP step, loads issued together [2]
 *   "batch_grid"   0..64    batched-closure workgroups per CU per launch [64]
 *   "batch_wait_us" >= 0    longest in-kernel wait of a fap_fw_batched_persistent task
 *                           for its matrix's earlier phases, in microseconds [2 s]
 *   "sssp_index"   0/1      batched part solves compute stage 1 through the graph's
 *                           boundary index (built by the first such call) [1]
 *   "bidx_tile"    0/64/128 columns of the boundary index's 64-row block-update tiles [128]
 *   "bidx_delta"   >= 0     distance band of the boundary index's block iterations; any
 *                           value gives the same rows [mean arc weight x sqrt(V / K),
 *                           times 32 / parts for calls of fewer than 32 parts] */
int fap_ctx_set_option(fap_ctx *ctx, const char *name, int64_t value);
int fap_ctx_get_option(fap_ctx *ctx, const char *name, int64_t *value);
int fap_ctx_sync(fap_ctx *ctx);
int fap_dev_alloc(fap_ctx *ctx, size_t bytes, void **dptr);
int fap_dev_free(fap_ctx *ctx, void *dptr);
int fap_memcpy_h2d(fap_ctx *ctx, void *dst, const void *src, size_t bytes);
int fap_memcpy_d2h(fap_ctx *ctx, void *dst, const void *src, size_t bytes);
int fap_memcpy_d2d(fap_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Fill n int32 with value (a kernel, not hipMemset; asynchronous). */
int fap_fill_i32(fap_ctx *ctx, int32_t *d, int64_t n, int32_t value, void *stream);
/* Device copy of n int32 by a streaming kernel (asynchronous). */
int fap_copy_i32(fap_ctx *ctx, int32_t *dst, const int32_t *src, int64_t n, void *stream);
/* Wrapping sum of n uint32 and of (index+1)*value (a position-keyed checksum);
 * synchronous. */
int fap_checksum_i32(fap_ctx *ctx, const int32_t *d, int64_t n, uint64_t *sum,
                     uint64_t *weighted, void *stream);
int fap_stream_create(fap_ctx *ctx, void **stream);
int fap_stream_destroy(fap_ctx *ctx, void *stream);
int fap_stream_sync(void *stream);
/* Events for timing: record on a stream, elapsed milliseconds between two. */
int fap_event_create(void **ev);
int fap_event_destroy(void *ev);
int fap_event_record(void *ev, void *stream);
int fap_event_elapsed(void *start, void *stop, float *ms);

/* ---- resident graph ---- */
/* Upload a CSR (host arrays: row_offset[V+1], col[E], w[E], int32, w >= 0).
 * Synchronous.  The graph stays in HBM until fap_graph_free. */
int fap_graph_upload(fap_ctx *ctx, int32_t V, int64_t E,
                     const int32_t *row_offset, const int32_t *col,
                     const int32_t *w, fap_graph_dev **out);
/* Attach the partition: graph_id (1-based parts), st2ed, ed2st (host, V
 * each), as produced by preCondition.  Synchronous. */
int fap_graph_set_partition(fap_ctx *ctx, fap_graph_dev *g,
                            const int32_t *graph_id, const int32_t *st2ed,
                            const int32_t *ed2st);
int fap_graph_free(fap_ctx *ctx, fap_graph_dev *g);
/* Near-far band width for fap_sssp_multi on this graph (0 = the context's
 * choice: option sssp_delta, the tune cache, else 8 / 6 / 4 x mean arc
 * weight by source count).  Affects speed only, never the distances. */
int fap_graph_set_delta(fap_graph_dev *g, int32_t delta);
int fap_graph_info(const fap_graph_dev *g, int32_t *V, int64_t *E);

/* ---- the three stages ---- */
/* Multi-source SSSP: row i of d_out (stride ld_out >= V) receives the
 * distances from d_sources[i] to every vertex (column = vertex id).
 * Strict relaxation from INF; rows are exact shortest-path distances. */
int fap_sssp_multi(fap_ctx *ctx, const fap_graph_dev *g,
                   const int32_t *d_sources, int32_t n_src, int32_t *d_out,
                   int64_t ld_out, void *stream);

/* In-place Floyd-Warshall closure of the n x n matrix at d_mat (stride ld),
 * blocked: diagonal tile in LDS, row/column panels, remaining tiles. */
int fap_fw_blocked(fap_ctx *ctx, int32_t *d_mat, int32_t n, int64_t ld,
                   void *stream);

/* Closure of `count` matrices (d_mats[j]: n[j] x n[j], stride ld[j]; host
 * arrays of device pointers and sizes) together: 64 x 64 tiles, one launch
 * per phase across all matrices (pivot rounds in pairs: diag, panels and
 * cross tiles of each round, one 128-deep pass over the remaining tiles of
 * both), each a persistent grid over a size-ordered task queue.  Same
 * results as fap_fw_blocked on each matrix.  Synchronous. */
int fap_fw_batched(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                   const int64_t *ld, void *stream);

/* fap_fw_batched without device allocation, for callers that cannot call
 * hipMalloc (code under stream capture, tight serving loops, harnesses that
 * forbid it).  No counterpart in the reference.  The batch's descriptors
 * (24 bytes per matrix + 16 per phase segment) go to a fixed device array
 * of FAP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a
 * batch that needs more is refused with FAP_E_ARG.  One chain of launches on
 * `stream`, same results as fap_fw_batched.  Calls that overlap on one GPU
 * (other threads, other streams) share that array and race; calls on one
 * stream are ordered.  Synchronous. */
#define FAP_NOALLOC_SCRATCH_BYTES (4 << 20)
int fap_fw_batched_noalloc(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                           const int64_t *ld, void *stream);

/* fap_fw_batched as ONE launch: the matrices are split over the device's
 * XCDs by work, each XCD's matrices go into one queue (their phase launches
 * laid end to end, larger matrices first within each), and the workgroups
 * on each XCD claim that queue's tasks in order from a counter; a task
 * waits, inside the kernel, until its own matrix's earlier phases have
 * completed.  Replaces the reference's host loop over parts
 * (fap/fap.cpp:189-193) and its per-process parts (fap/fap_more_task_mpi.cpp:34-91)
 * with one launch; same results as fap_fw_batched.  Every wait is bounded
 * (option "batch_wait_us"): a wait that runs out stops the launch and the
 * call returns FAP_E_DEVICE with the matrices partly closed (also if a
 * queue's XCD ran no workgroup).  Rows must be 128-byte aligned (each
 * matrix's address a multiple of 128 bytes, ld a multiple of 32), else
 * FAP_E_ARG: matrices on different XCDs then never share a cache line.
 * Slower than the plans' phase launches on the 512-graph batch (22.6 vs
 * 16.5 ms).  Synchronous. */
int fap_fw_batched_persistent(fap_ctx *ctx, int32_t count, int32_t *const *d_mats,
                              const int32_t *n, const int64_t *ld, void *stream);

/* A reusable batch: descriptors uploaded once; with use_graph the phase
 * launches are captured in a hipGraph and each run replays it (the
 * matrices are read and written in place at every run).  A plan of 8 or
 * more matrices runs them as two chains of launches, the second on a
 * stream the plan owns, forked from and joined back to the run's stream.
 * fap_fw_plan_info: launches summed over the chains, tasks, ops. */
typedef struct fap_fw_plan fap_fw_plan;
int fap_fw_plan_create(fap_ctx *ctx, int32_t count, int32_t *const *d_mats, const int32_t *n,
                       const int64_t *ld, int use_graph, fap_fw_plan **out);
/* A plan whose every run is ONE persistent launch (as
 * fap_fw_batched_persistent, same requirements): queues, descriptors and
 * the task map uploaded once; with use_graph each run replays a captured
 * hipGraph of a counter reset and the launch.  Runs are asynchronous: a
 * wait that gave up (or a queue left unfinished) is reported by
 * fap_fw_plan_check, which waits for the plan's last run. */
int fap_fw_plan_create_persistent(fap_ctx *ctx, int32_t count, int32_t *const *d_mats,
                                  const int32_t *n, const int64_t *ld, int use_graph,
                                  fap_fw_plan **out);
/* Waits for the plan's last run; FAP_E_DEVICE if it was a persistent run
 * that gave up or left a matrix unfinished, FAP_OK otherwise. */
int fap_fw_plan_check(fap_fw_plan *plan);
int fap_fw_plan_run(fap_fw_plan *plan, void *stream);
int fap_fw_plan_info(const fap_fw_plan *plan, int32_t *launches, int64_t *tasks, int64_t *ops);
int fap_fw_plan_destroy(fap_fw_plan *plan);

/* Dense distance matrix of each whole graph (K = 1 part): INF, then every
 * arc in CSR order (last duplicate wins), diagonal 0 -- what
 * MysubMatBuild_path builds when a part is the whole graph
 * (fap/utils/util_Myalgorithm.h:30-73).  d_mats[j] is V_j x V_j, stride
 * ld[j].  Two launches for the whole batch; synchronous. */
int fap_dense_from_graphs(fap_ctx *ctx, int32_t count, const fap_graph_dev *const *graphs,
                          int32_t *const *d_mats, const int64_t *ld, void *stream);

/* Min-plus product: C[i][j] = min(base, min_k A[i][k] + B[k][j]) where base
 * is C[i][j] if accumulate else INF.  A is M x K (lda), B is K x N (ldb),
 * C is M x N (ldc).  K may be 0 (C = base). */
int fap_minplus(fap_ctx *ctx, const int32_t *A, int64_t lda, const int32_t *B,
                int64_t ldb, int32_t *C, int64_t ldc, int32_t M, int32_t N,
                int32_t K, int accumulate, void *stream);

/* ---- part glue ---- */
/* Dense part matrix sub (n_sub x n_sub, stride ld_sub) for the part whose
 * vertices are st2ed[start .. start+n_sub), boundary first (n_bdy of them):
 * rows < n_bdy gathered from rowblock (stride ld_rb, columns = vertex ids),
 * other rows INF plus the part's internal arcs (last duplicate wins),
 * diagonal 0.  Needs fap_graph_set_partition. */
int fap_gather_submat(fap_ctx *ctx, const fap_graph_dev *g, int32_t start,
                      int32_t n_sub, int32_t n_bdy, const int32_t *rowblock,
                      int64_t ld_rb, int32_t *sub, int64_t ld_sub,
                      void *stream);
/* rowblock[i][st2ed[start+j]] = sub[i][j] for i < rows, j < n_sub. */
int fap_scatter_submat(fap_ctx *ctx, const fap_graph_dev *g, int32_t start,
                       int32_t rows, int32_t n_sub, const int32_t *sub,
                       int64_t ld_sub, int32_t *rowblock, int64_t ld_rb,
                       void *stream);

/* One part end to end, device resident: rowblock (n_sub x V, stride ld_rb)
 * receives the distances from every vertex of part `part_id` (1-based) in
 * st2ed order to all V vertices.  start/n_sub/n_bdy are the part's offset,
 * size and boundary count from preCondition.  Uses context workspace. */
int fap_subgraph_solve(fap_ctx *ctx, const fap_graph_dev *g, int32_t start,
                       int32_t n_sub, int32_t n_bdy, int32_t *rowblock,
                       int64_t ld_rb, void *stream);

/* Batched form of fap_subgraph_solve over `count` parts (host arrays of
 * starts, sizes, boundary counts and device row-block pointers, stride
 * ld_rb): every part's boundary rows through the graph's boundary index
 * (built by the first such call when "sssp_index" is on; one SSSP launch
 * for all boundary vertices when it is off or the index declines), one
 * batched closure of all part matrices, then per-part min-plus and scatter.
 * Same rows as fap_subgraph_solve on each part.  Synchronous. */
int fap_subgraph_solve_batch(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                             const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                             int32_t *const *rowblocks, int64_t ld_rb, void *stream);

/* Stage 1 alone for `count` parts (the reference's handle_boundry_path,
 * fap/kernel/batch_sssp.h:254-277, called once per part by run(),
 * fap/fap.cpp:105-110): rows 0..n_bdys[p]-1 of each part's row block
 * (ld_rb >= V) get the distances from the part's boundary vertices to every
 * vertex -- through the graph's boundary index (built by the first call
 * when the context's "sssp_index" is on), else one SSSP launch over all the
 * parts' boundary vertices.  The other rows are left as they are.
 * Synchronous. */
int fap_subgraph_boundary_rows(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                               const int32_t *starts, const int32_t *n_subs,
                               const int32_t *n_bdys, int32_t *const *rowblocks, int64_t ld_rb,
                               void *stream);
/* fap_subgraph_solve_batch over groups of `group_parts` parts, alternating
 * between the context's stream and a second context it creates on first
 * use: one group's SSSP (HBM-bound) runs beside the previous group's
 * closure and products (VALU-bound).  Same rows.  Synchronous. */
int fap_subgraph_solve_groups(fap_ctx *ctx, const fap_graph_dev *g, int32_t count,
                              const int32_t *starts, const int32_t *n_subs, const int32_t *n_bdys,
                              int32_t *const *rowblocks, int64_t ld_rb, int32_t group_parts);

/* Predecessor ("path") rows for exact distance rows -- the path matrices the
 * reference builds when is_path_needed (fap/kernel/batch_sssp.h:82-83,110,
 * fap/kernel/floyd.h:44, fap/kernel/minplus.h:57-60,
 * fap/utils/util_Myalgorithm.h:50-51,71): pred[i][j] is the vertex before j
 * on a shortest path from row i's vertex to j, the row's vertex itself at
 * j == that vertex, -1 when j is unreachable.  d_row_vertex holds the rows'
 * vertices (device).  Which of several equally short paths is not fixed.
 * Builds the graph's transposed CSR on first use.  Synchronous. */
int fap_pred_rows(fap_ctx *ctx, fap_graph_dev *g, const int32_t *d_dist, int64_t ld,
                  const int32_t *d_row_vertex, int32_t nrows, int32_t *d_pred, int64_t ld_pred,
                  void *stream);
/* The same for a part's row block (rows = st2ed[start .. start + n_sub)). */
int fap_subgraph_pred(fap_ctx *ctx, fap_graph_dev *g, int32_t start, int32_t n_sub,
                      const int32_t *rowblock, int64_t ld_rb, int32_t *predblock, int64_t ld_pred,
                      void *stream);

/* Boundary index of a partitioned graph: for every part, the distances
 * inside the part from its boundary vertices to its vertices (the closure of
 * the part's own arcs), plus the graph of boundary vertices (cut arcs and
 * those in-part distances).  Batched part solves then compute stage 1 -- the
 * SSSP rows from every boundary vertex, fap/kernel/batch_sssp.h:254-277 --
 * as an SSSP on that much smaller graph (its distance array in LDS) followed
 * by one min-plus expansion per part, exactly the same rows.  Built on the
 * first batched solve (option "sssp_index"); this call builds it at once.
 * *built = 0 when the graph does not qualify (more than ~36 k boundary
 * vertices, duplicate arcs, a partition whose parts do not list their
 * boundary vertices first); stage 1 then runs fap_sssp_multi.  The index is
 * dropped by fap_graph_set_partition and fap_graph_free.  Synchronous. */
int fap_graph_build_index(fap_ctx *ctx, fap_graph_dev *g, int32_t *built);
/* boundary vertices held by the index (0 without one) and its device bytes */
int fap_graph_index_info(const fap_graph_dev *g, int32_t *boundary, int64_t *bytes);

/* Min-plus product with the path matrix of min_plus_path_CPU_advanced
 * (fap/kernel/minplus.h:51-65): C = min(INF, A (x) B) and, when C_path is
 * not NULL, C_path[i][j] = B_path[k][j] for the first k attaining the
 * minimum (the reference's strict relaxation in k order), -1 where C is
 * INF -- bit-identical to the reference for any inputs. */
int fap_minplus_path(fap_ctx *ctx, const int32_t *A, int64_t lda, const int32_t *B, int64_t ldb,
                     const int32_t *B_path, int64_t ldbp, int32_t *C, int64_t ldc,
                     int32_t *C_path, int64_t ldcp, int32_t M, int32_t N, int32_t K,
                     void *stream);
/* Floyd-Warshall closure with the path matrix of floyd_CPU_path
 * (fap/kernel/floyd.h:38-49, path[i][j] = path[k][j] on each strict
 * improvement): d_path holds the caller's labels on entry.  A cell the
 * closure does not improve keeps its label; an improved cell takes the
 * input label of m -> j for a last hop of a shortest i -> j path in the
 * input matrix, chosen so that the hops lead back to i (which hop among
 * equal paths is not fixed: the reference's backends differ there).  Uses
 * 12 n^2 bytes of context workspace; synchronous. */
int fap_fw_blocked_path(fap_ctx *ctx, int32_t *d_mat, int32_t n, int64_t ld, int32_t *d_path,
                        int64_t ldp, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FAP_DEV_H_ */
