// A caller written against the reference's C++ API (fap/fap.h), used the way
// the reference's single-node example uses it (app/singleNodeExample.cpp:40-68):
// construct from a .mtx name, preCondition, solve part by part when the graph
// is split (with paths) or whole, read the current part back and verify it.
// Built by tests/test_cpp_drop_in.py against include/fap/fap.h and
// libfap_amd.so; prints one line "part <id> ok" or "part <id> wrong".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "fap/fap.h"

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s <graph name> <K> <directed 0|1> <weighted 0|1>\n", argv[0]);
    return 2;
  }
  const std::string name = argv[1];
  const int K = atoi(argv[2]);
  const bool directed = strcmp(argv[3], "1") == 0;
  const bool weighted = strcmp(argv[4], "1") == 0;

  fap::fapGraph G(name, directed, weighted, K);
  G.preCondition();
  int32_t rc = 0;
  if (G.isSplit() && K >= 3) {
    for (int id = 1; id <= 3; id++) rc |= G.solveSubGraph(id, true);
  } else {
    rc = G.solve();
  }
  if (rc != 0) {
    printf("solve failed: %s\n", G.lastError().c_str());
    return 1;
  }
  const int32_t id = G.getCurrentSubGraphId();
  std::vector<float> dist = G.getSubGraphDistance(id);
  std::vector<int32_t> path = G.getSubGraphPath(id);  // empty after solve() without paths
  std::vector<int32_t> source = G.getSubGraphIndex(id);
  const bool ok = G.check_result(dist.data(), path.empty() ? nullptr : path.data(), source.data(),
                                 (int)source.size(), (int)G.adj_size.size(), G.adj_size.data(),
                                 G.row_offset.data(), G.col_val.data(), G.weight.data(),
                                 G.getGraphId().data());
  // the path rows, when produced, walk back to their row's vertex
  size_t walks = 0;
  if (!path.empty()) {
    const int V = G.num_vertexs;
    for (size_t r = 0; r < source.size(); r += 97) {
      for (int t = 0; t < V; t += 131) {
        int x = t, steps = 0;
        while (x >= 0 && x != source[r] && steps <= V) x = path[r * V + x], steps++;
        if (x == source[r]) walks++;
        else if (dist[r * V + t] < 1e8f) {
          printf("part %d wrong path\n", id);
          return 1;
        }
      }
    }
  }
  printf("part %d %s (%zu path walks)\n", id, ok ? "ok" : "wrong", walks);
  return ok ? 0 : 1;
}
