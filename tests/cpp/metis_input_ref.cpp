// The xadj / adjncy that the reference's Graph_decomposition hands to
// METIS_PartGraphKway (fap/utils/readIdFile.h:71-112), restated: an
// undirected graph's CSR as it is; a directed graph symmetrised through a
// std::map<int, std::unordered_set<int>> (self loops skipped), vertices in
// map order and each vertex's neighbours in the set's iteration order.
// Reads "directed V E" then row_offset (V + 1) and col (E) from stdin and
// prints xadj and adjncy, one line each.
#include <stdio.h>

#include <map>
#include <unordered_set>
#include <vector>

int main() {
  int directed = 0, V = 0, E = 0;
  if (scanf("%d %d %d", &directed, &V, &E) != 3) return 2;
  std::vector<int> ro(V + 1), col(E);
  for (int &x : ro) scanf("%d", &x);
  for (int &x : col) scanf("%d", &x);
  std::vector<int> xadj(1, 0), adjncy;
  if (directed) {
    std::map<int, std::unordered_set<int>> graph;
    for (int ver = 0; ver < V; ver++)
      for (int e = ro[ver]; e < ro[ver + 1]; e++) {
        const int next = col[e];
        if (next == ver) continue;
        graph[ver].insert(next);
        graph[next].insert(ver);
      }
    for (auto &kv : graph) {
      for (int v : kv.second) adjncy.push_back(v);
      xadj.push_back(xadj.back() + (int)kv.second.size());
    }
  } else {
    xadj.assign(ro.begin(), ro.end());
    adjncy.assign(col.begin(), col.end());
  }
  for (int x : xadj) printf("%d ", x);
  printf("\n");
  for (int x : adjncy) printf("%d ", x);
  printf("\n");
  return 0;
}
