// fap/kernel.h -- the reference's operator names, forwarding to the MI355X
// C ABI (fap_ops.h).  Same signatures as fap/kernel/batch_sssp.h:254-257,
// fap/kernel/floyd.h:150 and fap/kernel/minplus.h:78-79 (namespace fap, void,
// host float pointers); a failure prints the reason and exits with
// EXIT_FAILURE, as the reference's HIP backend does on a HIP error.
#ifndef FAP_KERNEL_H_MI355X_
#define FAP_KERNEL_H_MI355X_

#include <stdio.h>
#include <stdlib.h>

#include "fap_ops.h"

namespace fap {

inline void mi355x_check(int rc, const char *what) {
  if (rc != FAP_OK) {
    fprintf(stderr, "error: %s: %s\n", what, fap_last_error());
    exit(EXIT_FAILURE);
  }
}

inline void handle_boundry_path(float *subGraph, int *subGraph_path, int vertexs, int edges,
                                int bdy_num, int *adj_size, int *row_offset, int *col_val,
                                float *weights, int *st2ed, int offset) {
  mi355x_check(fap_handle_boundry_path(subGraph, subGraph_path, vertexs, edges, bdy_num,
                                       adj_size, row_offset, col_val, weights, st2ed, offset),
               "handle_boundry_path");
}

inline void floyd_path(int num, float *mat, int *path) {
  mi355x_check(fap_floyd_path(num, mat, path), "floyd_path");
}

inline void min_plus_path_advanced(float *mat1, float *mat2, int *mat2_path, float *res,
                                   int *path, int M, int N, int K) {
  mi355x_check(fap_min_plus_path_advanced(mat1, mat2, mat2_path, res, path, M, N, K),
               "min_plus_path_advanced");
}

}  // namespace fap

#endif  // FAP_KERNEL_H_MI355X_
