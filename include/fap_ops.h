/*
 * fap_ops.h -- the reference's stage operators, host memory, float, C ABI.
 *
 * Same arguments and meaning as the three functions the reference's
 * fapGraph::run calls (host pointers owned by the caller, row-major, no
 * leading dimensions, float distances with INF = 1e8f), so a binding or a
 * WITH_MI355X branch in the reference's dispatch headers can forward to them
 * unchanged.  They run on device 0 through a process-wide context: copy in,
 * compute in int32, copy out.  Values >= 1e8 are read as INF; finite values
 * must be non-negative integers < 2^24 (the int32 build's exactness range).
 *
 * Replaces:
 *   fap_handle_boundry_path      <- fap::handle_boundry_path  fap/kernel/batch_sssp.h:254-277
 *   fap_floyd_path               <- fap::floyd_path           fap/kernel/floyd.h:150-158
 *   fap_min_plus_path_advanced   <- fap::min_plus_path_advanced fap/kernel/minplus.h:78-89
 *
 * Path arguments: may be NULL (no path computed).  When given they are
 * written with the reference's meaning: subGraph_path rows = Dijkstra
 * parents (fap/kernel/batch_sssp.h:82-83,110: the source itself at the
 * source, -1 where unreachable); floyd_path's path holds the caller's labels
 * on entry and each improved cell takes path[m][j] of a last hop m -> j of a
 * shortest path (fap/kernel/floyd.h:44); min_plus_path_advanced's path is
 * mat2_path[k][j] of the first minimising k, -1 where res is INF
 * (fap/kernel/minplus.h:57-60), exactly as the reference's CPU code.  Which
 * parent or hop is taken among equally short paths is not fixed (the
 * reference's CPU, CUDA and HIP backends differ there too).
 * Errors: a FAP_E_* code and fap_last_error(); the C++ wrappers in
 * fap/kernel.h print it and exit(EXIT_FAILURE), like the reference's HIP
 * backend (fap/kernel/hip/batch_sssp.cpp:13-22).
 */
#ifndef FAP_OPS_H_
#define FAP_OPS_H_

#include "fap_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rows i < bdy_num of subGraph (stride vertexs) = distances from
 * st2ed[offset + i]; adj_size may be NULL (row_offset carries the degrees) */
int fap_handle_boundry_path(float *subGraph, int *subGraph_path, int vertexs, int edges,
                            int bdy_num, int *adj_size, int *row_offset, int *col_val,
                            float *weights, int *st2ed, int offset);
/* in-place closure of the num x num matrix */
int fap_floyd_path(int num, float *mat, int *path);
/* res (M x N) = min(INF, mat1 (M x K) (x) mat2 (K x N)) */
int fap_min_plus_path_advanced(float *mat1, float *mat2, int *mat2_path, float *res, int *path,
                               int M, int N, int K);

#ifdef __cplusplus
}
#endif

#endif /* FAP_OPS_H_ */
