// Host-only exercise of the library's CPU code -- the .mtx reader, the
// partitioners (built-in bisection and the METIS adapter through a stand-in
// libmetis), findBoundry / sort_and_encode, the fapGraph C ABI up to
// preCondition and its getters, the tune-cache file format -- built with
// -fsanitize=address,undefined by tests/test_host.py.  Valid inputs must
// load, malformed ones must come back as error codes; ASan / UBSan abort on
// any memory or undefined-behaviour error on the way.  No GPU call is made.
// Usage: host_sanitize <scratch dir> [<stand-in libmetis>].  Prints "ok".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <random>
#include <string>
#include <vector>

#include "fap_dev.h"
#include "fap_graph.h"
#include "fap_tune.h"
#include "host.h"

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      printf("check failed at line %d: %s\n", __LINE__, #c);      \
      failures++;                                                 \
    }                                                             \
  } while (0)

static std::string g_dir;

static std::string write_file(const std::string &name, const std::string &text) {
  const std::string p = g_dir + "/" + name;
  FILE *f = fopen(p.c_str(), "w");
  if (!f) {
    printf("cannot write %s\n", p.c_str());
    exit(2);
  }
  fputs(text.c_str(), f);
  fclose(f);
  return p;
}

struct Arcs {
  int V;
  std::vector<int> u, v, w;
};

static Arcs random_graph(int V, int E, uint32_t seed, bool isolated_tail) {
  std::mt19937 rng(seed);
  Arcs a{V, {}, {}, {}};
  const int span = isolated_tail ? V - 3 : V;
  for (int e = 0; e < E; e++) {
    a.u.push_back((int)(rng() % span));
    a.v.push_back((int)(rng() % span));
    a.w.push_back(1 + (int)(rng() % 90));
  }
  return a;
}

static std::string mtx_text(const Arcs &a, bool weighted) {
  std::string s = "%%MatrixMarket matrix coordinate\n% comment\n";
  s += std::to_string(a.V) + " " + std::to_string(a.V) + " " + std::to_string(a.u.size()) + "\n";
  for (size_t e = 0; e < a.u.size(); e++) {
    s += std::to_string(a.u[e] + 1) + " " + std::to_string(a.v[e] + 1);
    if (weighted) s += " " + std::to_string(a.w[e]);
    s += "\n";
  }
  return s;
}

static void csr_checks(const fapi::HostCSR &g) {
  CHECK((int64_t)g.row_offset.size() == (int64_t)g.V + 1);
  CHECK((int64_t)g.col.size() == g.E && (int64_t)g.w.size() == g.E);
  CHECK(g.row_offset[0] == 0 && g.row_offset[g.V] == g.E);
  for (int64_t e = 0; e < g.E; e++) CHECK(g.col[e] >= 0 && g.col[e] < g.V);
}

static void reader() {
  const Arcs a = random_graph(200, 900, 7, true);
  for (int directed = 0; directed < 2; directed++)
    for (int weighted = 0; weighted < 2; weighted++) {
      const std::string p = write_file("g.mtx", mtx_text(a, weighted));
      fapi::HostCSR g;
      std::string err;
      CHECK(fapi::read_mtx(p, directed, weighted, &g, &err) == FAP_OK);
      CHECK(g.V == 200 && g.E == (int64_t)a.u.size() * (directed ? 1 : 2));
      csr_checks(g);
    }
  const char *bad[] = {
      "",                          // empty
      "3 3 2\n1 2 1.5\n2 3 2\n",   // non-integral weight
      "3 3 2\n1 2 1\n2 9 2\n",     // id out of range
      "3 3 2\n0 2 1\n",            // id 0 (1-based)
      "-4 -4 2\n1 2 1\n",          // negative size
      "3 3\n",                     // truncated header
      "2 2 1\n1 2 99999999999\n",  // weight beyond int32
      "3 3 2\n1 2 -1\n",           // negative weight is |c| in the reference
  };
  for (size_t i = 0; i < sizeof(bad) / sizeof(bad[0]); i++) {
    const std::string p = write_file("bad.mtx", bad[i]);
    fapi::HostCSR g;
    std::string err;
    const int rc = fapi::read_mtx(p, false, true, &g, &err);
    if (rc == FAP_OK) csr_checks(g);  // some are accepted as the reference reads them
  }
  fapi::HostCSR g;
  std::string err;
  CHECK(fapi::read_mtx(g_dir + "/does-not-exist.mtx", false, false, &g, &err) != FAP_OK);
}

static fapi::HostCSR csr_of(const Arcs &a, bool symmetric) {
  fapi::HostCSR g;
  g.V = a.V;
  std::vector<std::vector<std::pair<int, int>>> adj(a.V);
  for (size_t e = 0; e < a.u.size(); e++) {
    adj[a.u[e]].push_back({a.v[e], a.w[e]});
    if (symmetric) adj[a.v[e]].push_back({a.u[e], a.w[e]});
  }
  g.row_offset.assign(1, 0);
  for (auto &l : adj) {
    for (auto &x : l) {
      g.col.push_back(x.first);
      g.w.push_back(x.second);
    }
    g.row_offset.push_back((int32_t)g.col.size());
  }
  g.E = (int64_t)g.col.size();
  return g;
}

static void partitioners(const char *metis_lib) {
  const Arcs a = random_graph(500, 1500, 11, true);
  for (int sym = 0; sym < 2; sym++) {
    const fapi::HostCSR g = csr_of(a, sym);
    for (int K : {1, 2, 3, 7, 64, 500}) {
      std::vector<int32_t> gid;
      bool used = true;
      std::string err;
      setenv("FAP_NO_METIS", "1", 1);
      CHECK(fapi::partition_graph(g, !sym, K, &gid, &used, &err) == FAP_OK && !used);
      CHECK((int)gid.size() == g.V);
      for (int x : gid) CHECK(x >= 1 && x <= K);
      fapi::PartInfo pi;
      fapi::precondition(g, K, gid, &pi);
      CHECK((int)pi.st2ed.size() == g.V && (int)pi.ver_num.size() == K + 1);
    }
    std::vector<int32_t> gid;
    bool used = false;
    std::string err;
    CHECK(fapi::partition_graph(g, !sym, 0, &gid, &used, &err) != FAP_OK);
    CHECK(fapi::partition_graph(g, !sym, g.V + 1, &gid, &used, &err) != FAP_OK);
  }
  unsetenv("FAP_NO_METIS");
  if (metis_lib) {
    setenv("FAP_METIS_LIB", metis_lib, 1);
    setenv("FAP_METIS_RECORD", (g_dir + "/record.txt").c_str(), 1);
    const Arcs full = random_graph(300, 1200, 5, false);
    for (int sym = 0; sym < 2; sym++) {
      fapi::HostCSR g = csr_of(full, sym);
      std::vector<int32_t> gid;
      bool used = false;
      std::string err;
      const int rc = fapi::partition_graph(g, !sym, 9, &gid, &used, &err);
      // a directed random graph may leave a vertex without arcs: refused
      CHECK(rc == FAP_OK ? used && (int)gid.size() == g.V : !used);
    }
    const fapi::HostCSR iso = csr_of(a, false);  // isolated tail vertices
    std::vector<int32_t> gid;
    bool used = false;
    std::string err;
    CHECK(fapi::partition_graph(iso, true, 4, &gid, &used, &err) != FAP_OK);
    unsetenv("FAP_METIS_LIB");
  }
}

static void graph_api() {
  setenv("FAP_NO_METIS", "1", 1);
  const Arcs a = random_graph(400, 1600, 3, false);
  const fapi::HostCSR g = csr_of(a, true);
  std::vector<float> wf(g.w.begin(), g.w.end());
  fapg *h = nullptr;
  CHECK(fapg_create_from_csr(g.V, g.E, g.row_offset.data(), g.col.data(), wf.data(), 8, &h) ==
        FAP_OK);
  if (!h) return;
  std::vector<int32_t> gid(g.V), st(g.V), ed(g.V), flags(g.V);
  CHECK(fapg_get_graph_id(h, gid.data()) == FAP_OK);
  CHECK(fapg_precondition(h) == FAP_OK);
  std::vector<int32_t> vn(9), vo(9), bn(9), bo(9);
  CHECK(fapg_get_part_info(h, vn.data(), vo.data(), bn.data(), bo.data()) == FAP_OK);
  CHECK(fapg_get_mapping(h, st.data(), ed.data()) == FAP_OK);
  CHECK(fapg_get_boundary_flags(h, flags.data()) == FAP_OK);
  for (int i = 0; i < g.V; i++) CHECK(ed[st[i]] == i);
  std::vector<int32_t> idx(vn[1] > 0 ? vn[1] : 1);
  CHECK(fapg_get_subgraph_index(h, 1, idx.data()) == FAP_OK);
  CHECK(fapg_get_subgraph_index(h, 99, idx.data()) != FAP_OK);
  std::vector<int32_t> bad(g.V, 99);
  CHECK(fapg_set_partition(h, bad.data()) != FAP_OK);
  CHECK(fapg_destroy(h) == FAP_OK);
  // malformed CSRs
  std::vector<int32_t> ro = {0, 2, 1}, col = {1, 0};
  std::vector<float> w = {1.f, 1.f};
  fapg *x = nullptr;
  CHECK(fapg_create_from_csr(2, 2, ro.data(), col.data(), w.data(), 1, &x) != FAP_OK);
  ro = {0, 1, 2};
  col = {1, 7};
  CHECK(fapg_create_from_csr(2, 2, ro.data(), col.data(), w.data(), 1, &x) != FAP_OK);
  col = {1, 0};
  w = {0.5f, 1.f};
  CHECK(fapg_create_from_csr(2, 2, ro.data(), col.data(), w.data(), 1, &x) != FAP_OK);
  CHECK(fapg_create_from_csr(-1, 0, ro.data(), nullptr, nullptr, 1, &x) != FAP_OK);
  // from a file in the reference's directory layout
  const std::string sub = g_dir + "/graph";
  const std::string cmd = "mkdir -p " + sub + "/weight-directed";
  CHECK(system(cmd.c_str()) == 0);
  write_file("graph/weight-directed/t.mtx", mtx_text(a, true));
  if (chdir(g_dir.c_str()) == 0) {
    CHECK(fapg_create_from_file("t", 1, 1, 5, &x) == FAP_OK);
    if (x) {
      CHECK(fapg_precondition(x) == FAP_OK);
      fapg_destroy(x);
    }
    CHECK(fapg_create_from_file("missing", 1, 1, 5, &x) != FAP_OK);
  }
  unsetenv("FAP_NO_METIS");
}

static void tune_cache() {
  fap_tune_cache *c = nullptr;
  CHECK(fap_tune_cache_create(&c) == FAP_OK);
  const int32_t p[4] = {128, 1, 0, 0};
  CHECK(fap_tune_cache_set(c, "fw", 52, p, 2, 1.5) == FAP_OK);
  const std::string path = g_dir + "/tune.txt";
  CHECK(fap_tune_cache_save(c, path.c_str()) == FAP_OK);
  fap_tune_cache *d = nullptr;
  CHECK(fap_tune_cache_create(&d) == FAP_OK);
  CHECK(fap_tune_cache_load(d, path.c_str()) == FAP_OK);
  int32_t n = 0;
  CHECK(fap_tune_cache_size(d, &n) == FAP_OK && n == 1);
  write_file("tune_bad.txt", "garbage line\nfw x y\nfw 3 1 2 3 4 5 6 7 8 9 10 11 12\n\n#\n");
  (void)fap_tune_cache_load(d, (g_dir + "/tune_bad.txt").c_str());
  fap_tune_cache_destroy(c);
  fap_tune_cache_destroy(d);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  g_dir = argv[1];
  reader();
  partitioners(argc > 2 ? argv[2] : nullptr);
  graph_api();
  tune_cache();
  if (failures) {
    printf("%d checks failed\n", failures);
    return 1;
  }
  printf("ok\n");
  return 0;
}
