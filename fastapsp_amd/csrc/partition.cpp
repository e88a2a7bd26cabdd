// partition.cpp -- K-way partition on the host.
//
// The reference calls METIS_PartGraphKway with NULL options
// (fap/utils/readIdFile.h:71-132).  libmetis is loaded at run time when it is
// present (FAP_METIS_LIB or libmetis.so); the adapter builds the same xadj /
// adjncy the reference builds (CSR as is when undirected; the std::map of
// std::unordered_set symmetrisation when directed) and assumes METIS's
// default 32-bit idx_t / real_t.  When libmetis is absent the built-in
// recursive bisection below is used; distances do not depend on the
// partition, only the part sizes, boundary counts and the row order do.
#include <dlfcn.h>
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <string>
#include <unordered_set>
#include <vector>

#include "fap_dev.h"
#include "host.h"

namespace fapi {

namespace {

typedef int (*metis_kway_fn)(int32_t *, int32_t *, int32_t *, int32_t *, int32_t *, int32_t *,
                             int32_t *, int32_t *, float *, float *, int32_t *, int32_t *,
                             int32_t *);

metis_kway_fn load_metis() {
  const char *names[] = {getenv("FAP_METIS_LIB"), "libmetis.so", "libmetis.so.5"};
  for (const char *nm : names) {
    if (!nm || !*nm) continue;
    void *h = dlopen(nm, RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    void *f = dlsym(h, "METIS_PartGraphKway");
    if (f) return (metis_kway_fn)f;
    dlclose(h);
  }
  return nullptr;
}

int metis_partition(metis_kway_fn fn, const HostCSR &g, bool directed, int32_t K,
                    std::vector<int32_t> *gid, std::string *err) {
  std::vector<int32_t> xadj(1, 0), adjncy;
  if (directed) {
    std::map<int, std::unordered_set<int>> sym;
    for (int32_t u = 0; u < g.V; u++)
      for (int32_t e = g.row_offset[u]; e < g.row_offset[u + 1]; e++) {
        int32_t v = g.col[e];
        if (u == v) continue;
        sym[u].insert(v);
        sym[v].insert(u);
      }
    for (auto &kv : sym) {
      for (int v : kv.second) adjncy.push_back(v);
      xadj.push_back((int32_t)adjncy.size());
    }
    if ((int32_t)sym.size() != g.V) {
      *err = "METIS adapter: directed graph has vertices without arcs";
      return FAP_E_ARG;
    }
  } else {
    xadj.assign(g.row_offset.begin(), g.row_offset.end());
    adjncy.assign(g.col.begin(), g.col.end());
  }
  int32_t nv = g.V, ncon = 1, np = K, objval = 0;
  std::vector<int32_t> part(g.V, 0);
  int ret = fn(&nv, &ncon, xadj.data(), adjncy.data(), nullptr, nullptr, nullptr, &np, nullptr,
               nullptr, nullptr, &objval, part.data());
  if (ret != 1 /* METIS_OK */) {
    *err = "METIS_PartGraphKway failed with status " + std::to_string(ret);
    return FAP_E_ARG;
  }
  gid->resize(g.V);
  for (int32_t i = 0; i < g.V; i++) (*gid)[i] = part[i] + 1;
  return FAP_OK;
}

// Symmetric adjacency without self loops or duplicates.
struct SymGraph {
  std::vector<int32_t> off, adj;
};

SymGraph symmetrize(const HostCSR &g) {
  std::vector<std::pair<int32_t, int32_t>> arcs;
  arcs.reserve((size_t)g.E * 2);
  for (int32_t u = 0; u < g.V; u++)
    for (int32_t e = g.row_offset[u]; e < g.row_offset[u + 1]; e++) {
      int32_t v = g.col[e];
      if (u == v) continue;
      arcs.emplace_back(u, v);
      arcs.emplace_back(v, u);
    }
  std::sort(arcs.begin(), arcs.end());
  arcs.erase(std::unique(arcs.begin(), arcs.end()), arcs.end());
  SymGraph s;
  s.off.assign((size_t)g.V + 1, 0);
  for (auto &a : arcs) s.off[a.first + 1]++;
  for (int32_t v = 0; v < g.V; v++) s.off[v + 1] += s.off[v];
  s.adj.resize(arcs.size());
  for (size_t i = 0; i < arcs.size(); i++) s.adj[i] = arcs[i].second;
  return s;
}

class Bisector {
 public:
  Bisector(const SymGraph &g, int32_t V, std::vector<int32_t> *gid)
      : g_(g), region_(V, -1), dist_a_(V, -1), dist_b_(V, -1), gid_(gid) {}

  void run(int32_t K) {
    std::vector<int32_t> all(region_.size());
    std::iota(all.begin(), all.end(), 0);
    split(all, K, 1);
  }

 private:
  const SymGraph &g_;
  std::vector<int32_t> region_;
  std::vector<int32_t> dist_a_, dist_b_;
  std::vector<int32_t> *gid_;
  int32_t next_region_ = 0;

  // BFS restricted to region rid from src; fills dist (hops) for reached
  // vertices and returns them in visit order.
  void bfs(int32_t src, int32_t rid, std::vector<int32_t> &dist, std::vector<int32_t> &order) {
    order.clear();
    order.push_back(src);
    dist[src] = 0;
    for (size_t h = 0; h < order.size(); h++) {
      int32_t u = order[h];
      for (int32_t e = g_.off[u]; e < g_.off[u + 1]; e++) {
        int32_t v = g_.adj[e];
        if (region_[v] != rid || dist[v] >= 0) continue;
        dist[v] = dist[u] + 1;
        order.push_back(v);
      }
    }
  }

  static int32_t farthest(const std::vector<int32_t> &order, const std::vector<int32_t> &dist) {
    int32_t best = order[0];
    for (int32_t v : order)
      if (dist[v] > dist[best] || (dist[v] == dist[best] && v < best)) best = v;
    return best;
  }

  // Linear order of the region: per connected component (in order of the
  // smallest vertex id), vertices sorted by (da - db, da, id).
  std::vector<int32_t> linear_order(const std::vector<int32_t> &verts, int32_t rid) {
    std::vector<int32_t> out;
    out.reserve(verts.size());
    std::vector<int32_t> comp, tmp;
    for (int32_t v : verts) dist_a_[v] = -1;
    std::vector<int32_t> sorted(verts);
    std::sort(sorted.begin(), sorted.end());
    for (int32_t s : sorted) {
      if (dist_a_[s] >= 0) continue;  // already in an earlier component
      bfs(s, rid, dist_a_, comp);     // marks the component
      int32_t a = farthest(comp, dist_a_);
      for (int32_t v : comp) dist_a_[v] = -1;
      bfs(a, rid, dist_a_, tmp);
      int32_t b = farthest(tmp, dist_a_);
      for (int32_t v : comp) dist_b_[v] = -1;
      bfs(b, rid, dist_b_, tmp);
      std::vector<int32_t> c(comp);
      const std::vector<int32_t> &da = dist_a_, &db = dist_b_;
      std::sort(c.begin(), c.end(), [&](int32_t x, int32_t y) {
        int32_t kx = da[x] - db[x], ky = da[y] - db[y];
        if (kx != ky) return kx < ky;
        if (da[x] != da[y]) return da[x] < da[y];
        return x < y;
      });
      out.insert(out.end(), c.begin(), c.end());
      // keep dist_a_ >= 0 on this component so later seeds skip it
    }
    return out;
  }

  void split(const std::vector<int32_t> &verts, int32_t K, int32_t first_part) {
    if (K <= 1 || verts.size() <= 1) {
      for (int32_t v : verts) (*gid_)[v] = first_part;
      return;
    }
    const int32_t rid = next_region_++;
    for (int32_t v : verts) region_[v] = rid;
    std::vector<int32_t> ord = linear_order(verts, rid);
    const int32_t k1 = K / 2, k2 = K - k1;
    const size_t n1 = (size_t)((verts.size() * (uint64_t)k1 + K / 2) / (uint64_t)K);
    std::vector<int32_t> left(ord.begin(), ord.begin() + n1);
    std::vector<int32_t> right(ord.begin() + n1, ord.end());
    split(left, k1, first_part);
    split(right, k2, first_part + k1);
  }
};

}  // namespace

void bisection_partition(const HostCSR &g, int32_t K, std::vector<int32_t> *graph_id) {
  graph_id->assign(g.V, 1);
  if (K <= 1 || g.V == 0) return;
  SymGraph s = symmetrize(g);
  Bisector b(s, g.V, graph_id);
  b.run(K);
}

int partition_graph(const HostCSR &g, bool directed, int32_t K, std::vector<int32_t> *graph_id,
                    bool *used_metis, std::string *err) {
  *used_metis = false;
  if (K < 1 || K > std::max<int32_t>(g.V, 1)) {
    *err = "K must be in 1..V";
    return FAP_E_ARG;
  }
  if (K > 1 && !getenv("FAP_NO_METIS")) {
    metis_kway_fn fn = load_metis();
    if (fn) {
      int rc = metis_partition(fn, g, directed, K, graph_id, err);
      if (rc == FAP_OK) *used_metis = true;
      return rc;
    }
  }
  bisection_partition(g, K, graph_id);
  return FAP_OK;
}

}  // namespace fapi
