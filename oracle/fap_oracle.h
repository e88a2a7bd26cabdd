/*
 * fap_oracle.h -- CPU restatement of FastAPSP's subgraph-solve path in int32.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline.  The product (libfap_amd.so) never links or calls it.
 *
 * Arithmetic: int32 distances and weights, INF = ORC_INF = INT32_MAX/2.  The
 * reference computes in float with INF = 1e8f (fap/utils/parameter.h:26); the
 * two agree exactly while every finite distance is < 2^24 (float sums exact)
 * and < 1e8 (no saturation).  Every relaxation is strict (`>`), every
 * accumulator starts at INF, exactly as the cited reference lines.
 *
 * Each function names the reference lines it restates (paths relative to the
 * FastAPSP repository root).  How each is pinned (golden vectors from scipy,
 * closed forms, brute force, the reference's own CPU code built into
 * oracle/_ref) is stated in tests/test_oracle.py.
 */
#ifndef FAP_ORACLE_H_
#define FAP_ORACLE_H_

#include <stdint.h>

#define ORC_INF 0x3FFFFFFF

#ifdef __cplusplus
extern "C" {
#endif

/* Binary-heap Dijkstra from src over CSR (row_offset[V+1], col, w).
 * fap/kernel/batch_sssp.h:79-114 (dijkstra_path); path may be NULL. */
void orc_dijkstra(int32_t src, int32_t V, const int32_t *row_offset,
                  const int32_t *col, const int32_t *w, int32_t *dist,
                  int32_t *path);

/* Rows i = Dijkstra(sources[i]) written at out + i*ld, OpenMP over sources
 * like fap/kernel/batch_sssp.h:232-252 (batched_sssp_path), called from
 * handle_boundry_path :254-277 with sources = st2ed[offset..offset+b). */
void orc_batched_sssp(const int32_t *sources, int32_t S, int32_t V,
                      const int32_t *row_offset, const int32_t *col,
                      const int32_t *w, int32_t *out, int64_t ld,
                      int32_t *path_out, int nthreads);

/* In-place Floyd-Warshall, k-i-j order, strict relax.
 * fap/kernel/floyd.h:38-49 (floyd_CPU_path); path may be NULL. */
void orc_floyd(int32_t n, int32_t *mat, int32_t *path);

/* C[i][j] = min(INF, min_k A[i][k] + B[k][j]); path = B_path[k][j] at the
 * first strict improvement.  A is MxK, B is KxN, C is MxN, all row-major.
 * fap/kernel/minplus.h:51-65 (min_plus_path_CPU_advanced); paths may be NULL. */
void orc_min_plus(const int32_t *A, const int32_t *B, const int32_t *B_path,
                  int32_t *C, int32_t *C_path, int32_t M, int32_t N, int32_t K);

/* Accumulating variant C = min(C, A (x) B) used by the blocked Floyd-Warshall
 * phases (no reference counterpart as a free function; it is the update in
 * fap/kernel/floyd.h:87-104 without path). lda/ldb/ldc are row strides. */
void orc_min_plus_acc(const int32_t *A, int64_t lda, const int32_t *B,
                      int64_t ldb, int32_t *C, int64_t ldc,
                      int32_t M, int32_t N, int32_t K);

/* Dense n_sub x n_sub matrix of one part: intra-part CSR edges (last duplicate
 * wins), rows 0..bdy-1 gathered from the SSSP rows through st2ed, diagonal 0.
 * `sub` must be pre-filled with INF (and sub_path with -1) by the caller,
 * as fap/fap.cpp:101-103 does.  fap/utils/util_Myalgorithm.h:30-73. */
void orc_submat_build(int32_t *sub, int32_t *sub_path,
                      const int32_t *rowblock, const int32_t *rowblock_path,
                      const int32_t *graph_id, int32_t start, int32_t n_sub,
                      int32_t n_bdy, int32_t V, const int32_t *st2ed,
                      const int32_t *ed2st, const int32_t *row_offset,
                      const int32_t *col, const int32_t *w);

/* A[i][0..bdy) = sub[bdy+i][0..bdy).  fap/utils/util_Myalgorithm.h:88-95. */
void orc_mat1_build(int32_t *A, const int32_t *sub, int32_t inner,
                    int32_t bdy, int32_t n_sub);

/* rowblock[i][st2ed[start+j]] = sub[i][j].  util_Myalgorithm.h:75-86. */
void orc_submat_decode(const int32_t *sub, const int32_t *sub_path,
                       int32_t *rowblock, int32_t *rowblock_path,
                       int32_t start, int32_t rows, int32_t cols, int32_t V,
                       const int32_t *st2ed);

/* is_bdy[v] = 1 iff v is an endpoint of an arc whose endpoints lie in
 * different parts (part ids 1..K).  fap/utils/util_preprocess.h:32-58. */
void orc_find_boundary(int32_t K, int32_t V, const int32_t *graph_id,
                       const int32_t *row_offset, const int32_t *col,
                       int32_t *is_bdy);

/* Permutation sorted by (part asc, boundary first) with std::sort, and the
 * per-part counts/offsets (arrays of K+1, index 0 unused).
 * fap/utils/util_preprocess.h:60-106. */
void orc_sort_and_encode(int32_t K, int32_t V, const int32_t *graph_id,
                         const int32_t *is_bdy, int32_t *ver_num,
                         int32_t *ver_off, int32_t *bdy_num, int32_t *bdy_off,
                         int32_t *st2ed, int32_t *ed2st);

/* One part's row block (ver_num[id] x V, row-major) through the reference's
 * three stages: SSSP from the boundary, FW on the dense part matrix,
 * min-plus for the inner rows, FW block scattered back.
 * fap/fap.cpp:92-181 (fapGraph::run).  rowblock_path may be NULL. */
void orc_run_part(int32_t id, int32_t V, const int32_t *row_offset,
                  const int32_t *col, const int32_t *w,
                  const int32_t *graph_id, const int32_t *st2ed,
                  const int32_t *ed2st, const int32_t *ver_num,
                  const int32_t *ver_off, const int32_t *bdy_num,
                  int32_t *rowblock, int32_t *rowblock_path, int nthreads);

/* Validity of predecessor ("path") rows, the semantics of the reference's
 * path matrices (fap/kernel/batch_sssp.h:82-83,110; fap/kernel/floyd.h:44;
 * fap/kernel/minplus.h:57-60; fap/utils/util_Myalgorithm.h:50-51,71):
 * pred[i][own[i]] == own[i]; pred[i][j] == -1 iff dist[i][j] == INF;
 * otherwise an arc pred -> j with dist[i][pred] + w == dist[i][j] exists and
 * following predecessors from j reaches own[i] without a cycle.  Returns the
 * number of cells that break a rule (0 = valid).  Not a restatement: a
 * checker, since the reference's choice among equal paths differs between
 * its own backends. */
int64_t orc_check_pred(const int32_t *dist, int64_t ld, const int32_t *pred, int64_t ldp,
                       const int32_t *own, int32_t nrows, int32_t V, const int32_t *row_offset,
                       const int32_t *col, const int32_t *w);

#ifdef __cplusplus
}
#endif

#endif /* FAP_ORACLE_H_ */
