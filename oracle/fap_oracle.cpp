// fap_oracle.cpp -- CPU restatement (int32) of FastAPSP's subgraph-solve path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker or the CPU baseline.  The
// product library never links or calls it.
//
// Threading follows the reference: Dijkstra runs in parallel over sources
// (OpenMP, fap/kernel/batch_sssp.h:241-250); Floyd-Warshall and min-plus are
// single-threaded (fap/kernel/floyd.h:38-49, fap/kernel/minplus.h:51-65).

#include "fap_oracle.h"

#include <omp.h>

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

namespace {

inline int32_t deg(const int32_t *row_offset, int32_t v) {
  return row_offset[v + 1] - row_offset[v];
}

}  // namespace

extern "C" {

// fap/kernel/batch_sssp.h:79-114.  Lazy-deletion binary heap ordered by
// (distance, vertex); strict `>` relaxation; dist INF, path -1, src gets 0/src.
void orc_dijkstra(int32_t src, int32_t V, const int32_t *row_offset,
                  const int32_t *col, const int32_t *w, int32_t *dist,
                  int32_t *path) {
  std::fill(dist, dist + V, (int32_t)ORC_INF);
  if (path) std::fill(path, path + V, -1);
  dist[src] = 0;
  if (path) path[src] = src;
  std::vector<char> done(V, 0);
  typedef std::pair<int32_t, int32_t> Item;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  heap.push(Item(0, src));
  while (!heap.empty()) {
    Item top = heap.top();
    heap.pop();
    int32_t u = top.second;
    if (done[u]) continue;
    done[u] = 1;
    int32_t du = top.first;
    int32_t e0 = row_offset[u], n = deg(row_offset, u);
    for (int32_t i = 0; i < n; i++) {
      int32_t v = col[e0 + i];
      int32_t nd = du + w[e0 + i];
      if (dist[v] > nd) {
        dist[v] = nd;
        if (path) path[v] = u;
        heap.push(Item(nd, v));
      }
    }
  }
}

// fap/kernel/batch_sssp.h:232-252 (OpenMP over sources).
void orc_batched_sssp(const int32_t *sources, int32_t S, int32_t V,
                      const int32_t *row_offset, const int32_t *col,
                      const int32_t *w, int32_t *out, int64_t ld,
                      int32_t *path_out, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int32_t i = 0; i < S; i++) {
    orc_dijkstra(sources[i], V, row_offset, col, w, out + (int64_t)i * ld,
                 path_out ? path_out + (int64_t)i * ld : nullptr);
  }
}

// fap/kernel/floyd.h:38-49.
void orc_floyd(int32_t n, int32_t *mat, int32_t *path) {
  for (int64_t k = 0; k < n; k++) {
    for (int64_t i = 0; i < n; i++) {
      const int32_t dik = mat[i * n + k];
      int32_t *row = mat + i * n;
      const int32_t *krow = mat + k * n;
      for (int64_t j = 0; j < n; j++) {
        int32_t cand = dik + krow[j];
        if (row[j] > cand) {
          row[j] = cand;
          if (path) path[i * n + j] = path[k * n + j];
        }
      }
    }
  }
}

// fap/kernel/minplus.h:51-65.
void orc_min_plus(const int32_t *A, const int32_t *B, const int32_t *B_path,
                  int32_t *C, int32_t *C_path, int32_t M, int32_t N,
                  int32_t K) {
  for (int64_t i = 0; i < M; i++) {
    for (int64_t j = 0; j < N; j++) {
      int32_t best = ORC_INF;
      int32_t bp = -1;
      for (int64_t k = 0; k < K; k++) {
        int32_t cand = A[i * K + k] + B[k * N + j];
        if (best > cand) {
          best = cand;
          if (B_path) bp = B_path[k * N + j];
        }
      }
      C[i * N + j] = best;
      if (C_path) C_path[i * N + j] = bp;
    }
  }
}

void orc_min_plus_acc(const int32_t *A, int64_t lda, const int32_t *B,
                      int64_t ldb, int32_t *C, int64_t ldc, int32_t M,
                      int32_t N, int32_t K) {
  for (int64_t i = 0; i < M; i++) {
    for (int64_t k = 0; k < K; k++) {
      const int32_t a = A[i * lda + k];
      const int32_t *brow = B + k * ldb;
      int32_t *crow = C + i * ldc;
      for (int64_t j = 0; j < N; j++) {
        int32_t cand = a + brow[j];
        if (crow[j] > cand) crow[j] = cand;
      }
    }
  }
}

// fap/utils/util_Myalgorithm.h:30-73.
void orc_submat_build(int32_t *sub, int32_t *sub_path,
                      const int32_t *rowblock, const int32_t *rowblock_path,
                      const int32_t *graph_id, int32_t start, int32_t n_sub,
                      int32_t n_bdy, int32_t V, const int32_t *st2ed,
                      const int32_t *ed2st, const int32_t *row_offset,
                      const int32_t *col, const int32_t *w) {
  // intra-part arcs, CSR order, so the last duplicate arc wins
  for (int32_t i = 0; i < n_sub; i++) {
    int32_t u = st2ed[start + i];
    for (int32_t e = row_offset[u]; e < row_offset[u + 1]; e++) {
      int32_t v = col[e];
      if (graph_id[u] != graph_id[v]) continue;
      int32_t j = ed2st[v] - start;
      sub[(int64_t)i * n_sub + j] = w[e];
      if (sub_path) sub_path[(int64_t)i * n_sub + j] = u;
    }
  }
  // boundary rows come from the SSSP rows (columns are original ids)
  for (int32_t i = 0; i < n_bdy; i++) {
    for (int32_t j = 0; j < n_sub; j++) {
      int32_t v = st2ed[start + j];
      sub[(int64_t)i * n_sub + j] = rowblock[(int64_t)i * V + v];
      if (sub_path && rowblock_path)
        sub_path[(int64_t)i * n_sub + j] = rowblock_path[(int64_t)i * V + v];
    }
  }
  for (int32_t i = 0; i < n_sub; i++) {
    sub[(int64_t)i * n_sub + i] = 0;
    if (sub_path) sub_path[(int64_t)i * n_sub + i] = st2ed[start + i];
  }
}

// fap/utils/util_Myalgorithm.h:88-95.
void orc_mat1_build(int32_t *A, const int32_t *sub, int32_t inner,
                    int32_t bdy, int32_t n_sub) {
  for (int32_t i = 0; i < inner; i++)
    for (int32_t k = 0; k < bdy; k++)
      A[(int64_t)i * bdy + k] = sub[(int64_t)(i + bdy) * n_sub + k];
}

// fap/utils/util_Myalgorithm.h:75-86.
void orc_submat_decode(const int32_t *sub, const int32_t *sub_path,
                       int32_t *rowblock, int32_t *rowblock_path,
                       int32_t start, int32_t rows, int32_t cols, int32_t V,
                       const int32_t *st2ed) {
  for (int32_t i = 0; i < rows; i++) {
    for (int32_t j = 0; j < cols; j++) {
      int32_t v = st2ed[j + start];
      rowblock[(int64_t)i * V + v] = sub[(int64_t)i * cols + j];
      if (sub_path && rowblock_path)
        rowblock_path[(int64_t)i * V + v] = sub_path[(int64_t)i * cols + j];
    }
  }
}

// fap/utils/util_preprocess.h:32-58.  Only vertices whose part id lies in
// 1..K are scanned (the reference walks BlockVer[1..K]).
void orc_find_boundary(int32_t K, int32_t V, const int32_t *graph_id,
                       const int32_t *row_offset, const int32_t *col,
                       int32_t *is_bdy) {
  std::fill(is_bdy, is_bdy + V, 0);
  for (int32_t u = 0; u < V; u++) {
    if (graph_id[u] < 1 || graph_id[u] > K) continue;
    for (int32_t e = row_offset[u]; e < row_offset[u + 1]; e++) {
      int32_t v = col[e];
      if (graph_id[v] != graph_id[u]) {
        is_bdy[u] = 1;
        is_bdy[v] = 1;
      }
    }
  }
}

namespace {
struct OrcNode {
  int v;
  int part;
  int bdy;
};
bool orc_node_less(const OrcNode &a, const OrcNode &b) {
  if (a.part == b.part) return a.bdy > b.bdy;
  return a.part < b.part;
}
}  // namespace

// fap/utils/util_preprocess.h:60-106.  std::sort (unstable) on the sequence
// v = 0..V-1: the within-class order is the one libstdc++'s introsort yields.
void orc_sort_and_encode(int32_t K, int32_t V, const int32_t *graph_id,
                         const int32_t *is_bdy, int32_t *ver_num,
                         int32_t *ver_off, int32_t *bdy_num, int32_t *bdy_off,
                         int32_t *st2ed, int32_t *ed2st) {
  std::vector<OrcNode> nodes(V);
  for (int32_t i = 0; i < V; i++) nodes[i] = OrcNode{i, graph_id[i], is_bdy[i]};
  std::sort(nodes.begin(), nodes.end(), orc_node_less);
  for (int32_t i = 0; i < V; i++) {
    st2ed[i] = nodes[i].v;
    ed2st[nodes[i].v] = i;
  }
  std::vector<int32_t> nv(K + 1, 0), nb(K + 1, 0);
  for (int32_t i = 0; i < V; i++) {
    if (graph_id[i] >= 1 && graph_id[i] <= K) {
      nv[graph_id[i]]++;
      if (is_bdy[i]) nb[graph_id[i]]++;
    }
  }
  int32_t cnt = 0;
  ver_num[0] = ver_off[0] = bdy_num[0] = bdy_off[0] = 0;
  for (int32_t p = 1; p <= K; p++) {
    ver_off[p] = cnt;
    ver_num[p] = nv[p];
    bdy_off[p] = cnt;
    bdy_num[p] = nb[p];
    cnt += nv[p];
  }
}

// fap/fap.cpp:92-181 with the memory-split branch compiled out, as in the
// reference (WITH_GPU is never defined, :137).
void orc_run_part(int32_t id, int32_t V, const int32_t *row_offset,
                  const int32_t *col, const int32_t *w,
                  const int32_t *graph_id, const int32_t *st2ed,
                  const int32_t *ed2st, const int32_t *ver_num,
                  const int32_t *ver_off, const int32_t *bdy_num,
                  int32_t *rowblock, int32_t *rowblock_path, int nthreads) {
  const int32_t b = bdy_num[id];
  const int32_t n_sub = ver_num[id];
  const int32_t inner = n_sub - b;
  const int32_t start = ver_off[id];
  std::vector<int32_t> A((int64_t)inner * b, ORC_INF);
  std::vector<int32_t> sub((int64_t)n_sub * n_sub, ORC_INF);
  std::vector<int32_t> sub_path;
  if (rowblock_path) sub_path.assign((int64_t)n_sub * n_sub, -1);
  int32_t *sp = rowblock_path ? sub_path.data() : nullptr;

  // stage 1: SSSP from the part's boundary vertices into rows 0..b-1
  orc_batched_sssp(st2ed + start, b, V, row_offset, col, w, rowblock, V,
                   rowblock_path, nthreads);
  // stage 2: dense part matrix, Floyd-Warshall
  orc_submat_build(sub.data(), sp, rowblock, rowblock_path, graph_id, start,
                   n_sub, b, V, st2ed, ed2st, row_offset, col, w);
  orc_floyd(n_sub, sub.data(), sp);
  // stage 3: inner rows = (inner x b) (x) (b x V)
  orc_mat1_build(A.data(), sub.data(), inner, b, n_sub);
  int64_t off = (int64_t)b * V;
  orc_min_plus(A.data(), rowblock, rowblock_path, rowblock + off,
               rowblock_path ? rowblock_path + off : nullptr, inner, V, b);
  orc_submat_decode(sub.data(), sp, rowblock, rowblock_path, start, n_sub,
                    n_sub, V, st2ed);
}

int64_t orc_check_pred(const int32_t *dist, int64_t ld, const int32_t *pred, int64_t ldp,
                       const int32_t *own, int32_t nrows, int32_t V, const int32_t *row_offset,
                       const int32_t *col, const int32_t *w) {
  // in-arcs of every vertex: {source, weight}
  std::vector<int32_t> off(V + 1, 0);
  for (int32_t u = 0; u < V; u++)
    for (int32_t e = row_offset[u]; e < row_offset[u + 1]; e++) off[col[e] + 1]++;
  for (int32_t v = 0; v < V; v++) off[v + 1] += off[v];
  std::vector<int32_t> fill(off.begin(), off.end() - 1), isrc(off[V]), iw(off[V]);
  for (int32_t u = 0; u < V; u++)
    for (int32_t e = row_offset[u]; e < row_offset[u + 1]; e++) {
      isrc[fill[col[e]]] = u;
      iw[fill[col[e]]++] = w[e];
    }
  int64_t bad = 0;
  std::vector<int8_t> state(V);
  std::vector<int32_t> stack;
  for (int32_t i = 0; i < nrows; i++) {
    const int32_t *d = dist + (int64_t)i * ld;
    const int32_t *p = pred + (int64_t)i * ldp;
    const int32_t me = own[i];
    std::vector<int8_t> ok(V, 0);
    for (int32_t j = 0; j < V; j++) {
      if (j == me) {
        ok[j] = p[j] == me;
      } else if (d[j] >= ORC_INF) {
        ok[j] = p[j] == -1;
      } else if (p[j] >= 0 && p[j] < V && d[p[j]] < ORC_INF) {
        for (int32_t e = off[j]; e < off[j + 1] && !ok[j]; e++)
          ok[j] = isrc[e] == p[j] && d[p[j]] + iw[e] == d[j];
      }
      if (!ok[j]) bad++;
    }
    // walks: 0 unvisited, 1 on the current walk, 2 reaches me
    std::fill(state.begin(), state.end(), 0);
    if (me >= 0 && me < V) state[me] = 2;
    for (int32_t j = 0; j < V; j++) {
      if (state[j] || d[j] >= ORC_INF || !ok[j]) continue;
      stack.clear();
      int32_t x = j;
      while (x >= 0 && x < V && state[x] == 0 && ok[x] && d[x] < ORC_INF) {
        state[x] = 1;
        stack.push_back(x);
        x = p[x];
      }
      const bool reaches = x >= 0 && x < V && state[x] == 2;
      for (int32_t y : stack) state[y] = reaches ? 2 : 3;
      if (!reaches) bad += (int64_t)stack.size();
    }
  }
  return bad;
}

}  // extern "C"
