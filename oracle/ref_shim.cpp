// ref_shim.cpp -- C-ABI shim over the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference's headers where they lie (fap/kernel/*.h, fap/utils/*.h under the
// FastAPSP checkout); nothing of the reference is copied here.  The result,
// oracle/_ref/libfapref.so, pins the int32 restatement in fap_oracle.cpp and
// serves as the "reference" CPU baseline.  The reference's fap/fap.cpp cannot
// be compiled here (it needs metis.h, absent from this image), so
// ref_run_part below re-states the glue of fapGraph::run (fap/fap.cpp:92-181)
// and calls only the reference's own stage functions.
//
// Everything is float with INF = 1e8f, as in the reference.

#include <assert.h>
#include <string.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "fap/kernel/batch_sssp.h"
#include "fap/kernel/floyd.h"
#include "fap/kernel/minplus.h"
#include "fap/utils/readMatFile.h"
#include "fap/utils/util_Myalgorithm.h"
#include "fap/utils/util_preprocess.h"

extern "C" {

float ref_maxvalue() { return fap::MAXVALUE; }

// fap/utils/readMatFile.h:36-64 -- header counts (m doubled if undirected).
int ref_read_ver_edges(const char *name, int directed, int weighted, int *n,
                       int *m) {
  fap::readVerEdges(*n, *m, std::string(name), directed != 0, weighted != 0);
  return 0;
}

// fap/utils/readMatFile.h:66-173 -- CSR in the reference's adjacency order.
int ref_read_mat_file(const char *name, int directed, int weighted, int n,
                      int m, int *adj_size, int *row_offset, int *col_val,
                      float *weight) {
  fap::readMatFile(n, m, adj_size, row_offset, col_val, weight,
                   std::string(name), directed != 0, weighted != 0);
  return 0;
}

// fap/kernel/batch_sssp.h:79-114.
void ref_dijkstra_path(int u, int V, int *adj_size, int *row_offset,
                       int *col_val, float *weight, float *dist, int *path) {
  fap::dijkstra_path(u, V, adj_size, row_offset, col_val, weight, dist, path);
}

// fap/kernel/batch_sssp.h:254-277 (CPU branch, OpenMP over sources).
void ref_handle_boundry_path(float *sub_graph, int *sub_graph_path, int V,
                             int E, int bdy_num, int *adj_size,
                             int *row_offset, int *col_val, float *weight,
                             int *st2ed, int offset) {
  fap::handle_boundry_path(sub_graph, sub_graph_path, V, E, bdy_num, adj_size,
                           row_offset, col_val, weight, st2ed, offset);
}

// fap/kernel/floyd.h:38-49.
void ref_floyd_cpu_path(int n, float *mat, int *path) {
  fap::floyd_CPU_path(n, mat, path);
}

// fap/kernel/minplus.h:51-65 (called directly: the dispatcher at :78-89
// prints a debug line on every call).
void ref_min_plus_cpu_advanced(float *A, float *B, int *B_path, float *C,
                               int *C_path, int M, int N, int K) {
  fap::min_plus_path_CPU_advanced(A, B, B_path, C, C_path, M, N, K);
}

// fap/utils/util_Myalgorithm.h:30-73.
void ref_submat_build(float *sub, int *sub_path, float *rowblock,
                      int *rowblock_path, int *graph_id, int start, int n_sub,
                      int n_bdy, int V, int *st2ed, int *ed2st, int *adj_size,
                      int *row_offset, int *col_val, float *weight) {
  fap::MysubMatBuild_path(sub, sub_path, rowblock, rowblock_path, graph_id,
                          start, n_sub, n_bdy, V, st2ed, ed2st, adj_size,
                          row_offset, col_val, weight);
}

// fap/utils/util_Myalgorithm.h:88-95.
void ref_mat1_build(float *A, float *sub, int inner, int bdy, int n_sub) {
  fap::MyMat1Build(A, sub, inner, bdy, n_sub);
}

// fap/utils/util_Myalgorithm.h:75-86.
void ref_submat_decode(float *sub, int *sub_path, float *rowblock,
                       int *rowblock_path, int start, int rows, int cols,
                       int V, int *st2ed) {
  fap::MysubMatDecode_path(sub, sub_path, rowblock, rowblock_path, start, rows,
                           cols, V, st2ed);
}

// fap/utils/util_preprocess.h:32-106, with BlockVer filled the way
// Graph_decomposition does (fap/utils/readIdFile.h:127-131): vertex order.
void ref_precondition(int K, int V, int *graph_id, int *adj_size,
                      int *row_offset, int *col_val, float *weight,
                      int *is_bdy, int *ver_num, int *ver_off, int *bdy_num,
                      int *bdy_off, int *st2ed, int *ed2st) {
  std::unordered_map<int, std::vector<int>> block_ver, block_bdy;
  std::unordered_map<int, int> bdy;
  for (int i = 0; i < V; i++) block_ver[graph_id[i]].push_back(i);
  fap::findBoundry(K, V, graph_id, adj_size, row_offset, col_val, weight,
                   block_ver, block_bdy, bdy);
  for (int i = 0; i < V; i++) {
    auto it = bdy.find(i);
    is_bdy[i] = (it == bdy.end()) ? 0 : it->second;
  }
  fap::sort_and_encode(K, V, graph_id, bdy, ver_num, ver_off, bdy_num, bdy_off,
                       block_ver, block_bdy, st2ed, ed2st);
}

// Glue of fap/fap.cpp:92-181, calling the reference's stage functions.
void ref_run_part(int id, int V, int E, int *adj_size, int *row_offset,
                  int *col_val, float *weight, int *graph_id, int *st2ed,
                  int *ed2st, int *ver_num, int *ver_off, int *bdy_num,
                  float *rowblock, int *rowblock_path) {
  const int b = bdy_num[id];
  const int n_sub = ver_num[id];
  const int inner = n_sub - b;
  std::vector<float> A((int64_t)inner * b, fap::MAXVALUE);
  std::vector<float> sub((int64_t)n_sub * n_sub, fap::MAXVALUE);
  std::vector<int> sub_path((int64_t)n_sub * n_sub, -1);
  fap::handle_boundry_path(rowblock, rowblock_path, V, E, b, adj_size,
                           row_offset, col_val, weight, st2ed, ver_off[id]);
  fap::MysubMatBuild_path(sub.data(), sub_path.data(), rowblock, rowblock_path,
                          graph_id, ver_off[id], n_sub, b, V, st2ed, ed2st,
                          adj_size, row_offset, col_val, weight);
  fap::floyd_CPU_path(n_sub, sub.data(), sub_path.data());
  fap::MyMat1Build(A.data(), sub.data(), inner, b, n_sub);
  int64_t off = (int64_t)b * V;
  fap::min_plus_path_CPU_advanced(A.data(), rowblock, rowblock_path,
                                  rowblock + off, rowblock_path + off, inner,
                                  V, b);
  fap::MysubMatDecode_path(sub.data(), sub_path.data(), rowblock,
                           rowblock_path, ver_off[id], n_sub, n_sub, V,
                           st2ed);
}

}  // extern "C"
