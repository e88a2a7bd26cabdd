/*
 * fap_graph.h -- C ABI over fap::fapGraph (fap/fap.h), for FFI callers.
 *
 * One handle per graph.  Every function returns FAP_OK (0) or a negative
 * FAP_E_* code (fap_dev.h) and leaves the reason in fap_last_error().
 * Output buffers are host memory sized by the caller as documented.
 *
 * Reference interfaces mirrored (fap/fap.h and fap/fap.cpp of FastAPSP):
 *   fapg_create_from_file   <- fapGraph(file, directed, weighted, K)  fap/fap.cpp:27-59
 *   fapg_create_from_csr    <- fapGraph(V, E, row_offset, col, weight, K)  fap/fap.h:81-83
 *                              (a stub in the reference, fap/fap.cpp:61-65)
 *   fapg_precondition       <- preCondition()      fap/fap.cpp:67-81
 *   fapg_is_split           <- isSplit()           fap/fap.cpp:83-85
 *   fapg_get_graph_id       <- getGraphId()        fap/fap.cpp:87-89
 *   fapg_solve              <- solve()             fap/fap.cpp:184-194
 *   fapg_solve_subgraph     <- solveSubGraph(id)   fap/fap.cpp:197-207
 *   fapg_get_mapping        <- getMapping()        fap/fap.cpp:209-211
 *   fapg_get_all_distance_* <- getAllDistance()    fap/fap.cpp:213-216
 *   fapg_get_current_subgraph_id <- getCurrentSubGraphId() fap/fap.cpp:231-233
 *   fapg_get_subgraph_index <- getSubGraphIndex(id) fap/fap.cpp:223-229
 *   fapg_get_subgraph_distance_* <- getSubGraphDistance(id) fap/fap.cpp:235-238
 *   fapg_get_distance_from_one_point <- getDistanceFromOnePoint(v) fap/fap.cpp:245-252
 *   fapg_distance_p2p       <- getDistanceP2P()    fap/fap.cpp:263-265 (a stub there)
 *   fapg_check_result       <- check_result(...)   fap/fap.cpp:271-277
 *   fapg_get_all_path       <- getAllPath()        fap/fap.cpp:218-221
 *   fapg_get_subgraph_path  <- getSubGraphPath(id) fap/fap.cpp:240-243
 *   fapg_get_path_from_one_point <- getPathFromOnePoint(v) fap/fap.cpp:254-261
 *   fapg_path_p2p           <- getPathP2P()        fap/fap.cpp:267-269 (a stub there)
 */
#ifndef FAP_GRAPH_H_
#define FAP_GRAPH_H_

#include <stdint.h>

#include "fap_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fapg fapg;

int fapg_create_from_file(const char *name, int directed, int weighted, int32_t K, fapg **out);
/* weights as float like the reference constructor (integral values only) */
int fapg_create_from_csr(int32_t V, int64_t E, const int32_t *row_offset, const int32_t *col,
                         const float *weight, int32_t K, fapg **out);
int fapg_destroy(fapg *g);
/* 1-based part ids, V entries; replaces the constructor's partition */
int fapg_set_partition(fapg *g, const int32_t *graph_id);
int fapg_precondition(fapg *g);
int fapg_is_split(fapg *g, int *out);
int fapg_solve(fapg *g, int is_path_needed);
int fapg_solve_subgraph(fapg *g, int32_t id, int is_path_needed);
/* batched (1, default) or plain (0) part loop for solve / solve_subgraphs */
int fapg_set_batched(fapg *g, int batched);
/* several parts at once; each id's rows stay available to the getters */
int fapg_solve_subgraphs(fapg *g, int32_t count, const int32_t *ids, int is_path_needed);

int fapg_num_vertices(fapg *g, int32_t *V);
int fapg_num_edges(fapg *g, int64_t *E);
int fapg_num_parts(fapg *g, int32_t *K);
int fapg_is_directed(fapg *g, int *directed);
int fapg_used_metis(fapg *g, int *used);
/* CSR as stored (row_offset[V+1], col[E], int32 weights[E]) */
int fapg_get_csr(fapg *g, int32_t *row_offset, int32_t *col, int32_t *w);
int fapg_get_graph_id(fapg *g, int32_t *out /* V */);
int fapg_get_boundary_flags(fapg *g, int32_t *out /* V */);
int fapg_get_mapping(fapg *g, int32_t *st2ed /* V */, int32_t *ed2st /* V, may be NULL */);
/* per-part tables, K+1 entries each (index 0 unused, as the reference) */
int fapg_get_part_info(fapg *g, int32_t *ver_num, int32_t *ver_off, int32_t *bdy_num,
                       int32_t *bdy_off);
int fapg_get_current_subgraph_id(fapg *g, int32_t *id);
int fapg_get_subgraph_index(fapg *g, int32_t id, int32_t *out /* ver_num[id] */);
/* rows of part id (ver_num[id] x V): int32 (INF = FAP_INF) or float (INF = 1e8f) */
int fapg_get_subgraph_distance_i32(fapg *g, int32_t id, int32_t *out);
int fapg_get_subgraph_distance_f32(fapg *g, int32_t id, float *out);
/* V x V after fapg_solve */
int fapg_get_all_distance_i32(fapg *g, int32_t *out);
int fapg_get_all_distance_f32(fapg *g, float *out);
/* rows [row0, row0 + nrows) of a solved part (st2ed order), nrows x V int32;
 * no counterpart in the reference (its getters copy whole blocks) */
int fapg_get_subgraph_rows_i32(fapg *g, int32_t id, int32_t row0, int32_t nrows, int32_t *out);
/* row of vertex v in the current part (V floats) */
int fapg_get_distance_from_one_point(fapg *g, int32_t v, float *out);
int fapg_distance_p2p(fapg *g, int32_t src, int32_t dst, int32_t *out);
/* path rows (solved with is_path_needed): the vertex before each column on a
 * shortest path from the row's vertex, the vertex itself on its own column,
 * -1 when unreachable; shapes as the distance getters */
int fapg_get_subgraph_path(fapg *g, int32_t id, int32_t *out);
int fapg_get_all_path(fapg *g, int32_t *out);
int fapg_get_path_from_one_point(fapg *g, int32_t v, int32_t *out);
/* vertices src .. dst of a shortest path into out[cap]; *len = 0 when dst is
 * unreachable; FAP_E_ARG when cap is too small (*len says how many) */
int fapg_path_p2p(fapg *g, int32_t src, int32_t dst, int32_t *out, int32_t cap, int32_t *len);
/* device-side checksums of a solved part block (see fap_checksum_i32) */
int fapg_subgraph_checksum(fapg *g, int32_t id, uint64_t *sum, uint64_t *weighted);
/* check_ans restated (deterministic samples); *ok = 1 when all samples match */
int fapg_check_result(fapg *g, int32_t id, int *ok);

#ifdef __cplusplus
}
#endif

#endif /* FAP_GRAPH_H_ */
