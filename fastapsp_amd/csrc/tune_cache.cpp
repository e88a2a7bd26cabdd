// tune_cache.cpp -- the autotuner's cache (include/fap_tune.h): host only.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>

#include "common.h"
#include "fap_tune.h"

struct fap_tune_cache {
  struct Entry {
    int32_t p[2] = {0, 0};
    double ms = 0;
  };
  std::mutex mu;
  std::map<std::pair<std::string, int64_t>, Entry> entries;
};

namespace fapi {
namespace {

int nparams_of(const std::string &kind) {
  if (kind == "fw" || kind == "sssp") return 2;
  if (kind == "batch") return 1;
  return -1;
}

bool valid(const std::string &kind, const int32_t *p) {
  if (kind == "fw") return (p[0] == 64 || p[0] == 128) && (p[1] == 0 || p[1] == 1);
  if (kind == "sssp") return (p[0] == 256 || p[0] == 512 || p[0] == 1024) && p[1] >= 1 && p[1] <= 1024;
  if (kind == "batch") return p[0] >= 1 && p[0] <= 64;
  return false;
}

int64_t ceil_log2(uint64_t x) {
  int64_t k = 0;
  while (((uint64_t)1 << k) < x && k < 63) k++;
  return k;
}

fap_tune_cache *g_env_cache = nullptr;
std::once_flag g_env_once;

}  // namespace

bool tune_lookup(const fap_ctx *ctx, const char *kind, int64_t key, int32_t *params) {
  fap_tune_cache *c = ctx ? ctx->tune : nullptr;
  if (!c) return false;
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->entries.find({kind, key});
  if (it == c->entries.end()) return false;
  const int np = nparams_of(kind);
  for (int i = 0; i < np; i++) params[i] = it->second.p[i];
  return true;
}

fap_tune_cache *tune_env_cache() {
  std::call_once(g_env_once, [] {
    const char *path = getenv("FAP_TUNE_CACHE");
    if (!path || !*path) return;
    fap_tune_cache *c = new fap_tune_cache();
    if (fap_tune_cache_load(c, path) == FAP_OK) g_env_cache = c;
    else delete c;
  });
  return g_env_cache;
}

void fw_settings(const fap_ctx *ctx, int32_t n, int *tile, bool *lookahead) {
  int32_t p[2] = {64, 0};
  const bool cached = tune_lookup(ctx, "fw", fap_tune_key_fw(n), p);
  // default: 128 once phase 3 has at least 8 tiles of 128 x 128 per CU
  // (n >= ~6,000 on 256 CUs), where it measured faster; 64 below
  const int64_t nb128 = ((int64_t)n + 127) / 128;
  const int def_tile =
      (nb128 - 1) * (nb128 - 1) >= 8 * (int64_t)(ctx ? ctx->num_cus : 256) ? 128 : 64;
  *tile = (ctx && (ctx->opt_fw_tile == 64 || ctx->opt_fw_tile == 128)) ? ctx->opt_fw_tile
          : cached                                                     ? p[0]
                                                                       : def_tile;
  *lookahead = (ctx && ctx->opt_fw_lookahead >= 0) ? ctx->opt_fw_lookahead != 0
               : cached                            ? p[1] != 0
                                                   : false;
}

void sssp_settings(const fap_ctx *ctx, const fap_graph_dev *g, int32_t n_src, int *threads,
                   int *delta, int *group) {
  int32_t p[2] = {1024, 8};
  const bool cached = tune_lookup(ctx, "sssp", fap_tune_key_sssp(g->V, g->E, n_src), p);
  // Without a tuned entry: one 1024-thread workgroup per CU while the
  // sources fit on the CUs, smaller workgroups (more per CU) beyond, and a
  // band that narrows as more sources run at once (fewer re-relaxations
  // where the memory system is the limit, fewer rounds where it is not).
  // Measured on the 512x512 grid: 74 sources 10.2 ms at 1024 threads vs
  // 18.2 ms at 256, band 8 x mean weight best (10.9 ms at 4 x); 3669
  // sources 119 ms at 256 threads vs 165 ms at 1024, band 3-4 x best
  // (110 ms, 120 ms at 8 x).
  const int cus = ctx ? ctx->num_cus : 256;
  const int by_count = n_src <= cus ? 1024 : n_src <= 2 * cus ? 512 : 256;
  const double mult_by_count = n_src <= cus ? 8.0 : n_src <= 2 * cus ? 6.0 : 4.0;
  *threads = (ctx && ctx->opt_sssp_threads > 0) ? ctx->opt_sssp_threads : cached ? p[0] : by_count;
  // Grouped sources only on request: on the same grid a group of 8 takes
  // ~10x one source (its frontiers already keep a workgroup's lanes busy),
  // so grouping gains nothing there.
  *group = (ctx && ctx->opt_sssp_group > 0) ? ctx->opt_sssp_group : 1;
  if (ctx && ctx->opt_sssp_delta > 0) {
    *delta = ctx->opt_sssp_delta;
  } else if (g->delta > 0) {
    *delta = g->delta;
  } else if (g->E == 0) {
    *delta = 1;
  } else {
    const double mean = (double)g->weight_sum / (double)g->E;
    const double mult = cached ? (double)p[1] : mult_by_count;
    *delta = (int)std::max(1.0, std::min(1.0e6, mult * mean));
  }
}

int batch_grid_per_cu(const fap_ctx *ctx, int64_t batch_key) {
  if (ctx && ctx->opt_batch_grid > 0) return ctx->opt_batch_grid;
  // 64: 8 times the resident workgroups of the phase kernels (8 per CU) --
  // for most launches one task per workgroup, so the hardware hands out
  // tasks as workgroups finish (512 uneven graphs: 18.1 ms vs 19.0 at 16)
  int32_t p[1] = {64};
  if (tune_lookup(ctx, "batch", batch_key, p)) return p[0];
  return 64;
}

}  // namespace fapi

using namespace fapi;

extern "C" {

int fap_tune_cache_create(fap_tune_cache **out) {
  if (!out) return fail(FAP_E_ARG, "null output");
  *out = new fap_tune_cache();
  return FAP_OK;
}

int fap_tune_cache_destroy(fap_tune_cache *c) {
  delete c;
  return FAP_OK;
}

int fap_tune_cache_set(fap_tune_cache *c, const char *kind, int64_t key, const int32_t *params,
                       int32_t nparams, double ms) {
  if (!c || !kind || !params) return fail(FAP_E_ARG, "null argument");
  const std::string k(kind);
  const int np = nparams_of(k);
  if (np < 0) return fail(FAP_E_ARG, "unknown tune kind '" + k + "'");
  if (nparams != np) return fail(FAP_E_ARG, "wrong parameter count for '" + k + "'");
  if (!valid(k, params)) return fail(FAP_E_ARG, "parameter out of range for '" + k + "'");
  if (key < 0 || !(ms >= 0)) return fail(FAP_E_ARG, "bad key or time");
  std::lock_guard<std::mutex> lk(c->mu);
  fap_tune_cache::Entry &e = c->entries[{k, key}];
  for (int i = 0; i < np; i++) e.p[i] = params[i];
  e.ms = ms;
  return FAP_OK;
}

int fap_tune_cache_get(fap_tune_cache *c, const char *kind, int64_t key, int32_t *params,
                       int32_t nparams, double *ms) {
  if (!c || !kind || !params) return fail(FAP_E_ARG, "null argument");
  const std::string k(kind);
  const int np = nparams_of(k);
  if (np < 0 || nparams != np) return fail(FAP_E_ARG, "unknown kind or parameter count");
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->entries.find({k, key});
  if (it == c->entries.end()) return fail(FAP_E_STATE, "no entry");
  for (int i = 0; i < np; i++) params[i] = it->second.p[i];
  if (ms) *ms = it->second.ms;
  return FAP_OK;
}

int fap_tune_cache_size(fap_tune_cache *c, int32_t *n) {
  if (!c || !n) return fail(FAP_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *n = (int32_t)c->entries.size();
  return FAP_OK;
}

int fap_tune_cache_save(fap_tune_cache *c, const char *path) {
  if (!c || !path) return fail(FAP_E_ARG, "null argument");
  std::ostringstream os;
  os << "# fap tune cache: kind key params... ms\n";
  {
    std::lock_guard<std::mutex> lk(c->mu);
    for (const auto &kv : c->entries) {
      const int np = nparams_of(kv.first.first);
      os << kv.first.first << ' ' << kv.first.second;
      for (int i = 0; i < np; i++) os << ' ' << kv.second.p[i];
      char ms[32];
      snprintf(ms, sizeof(ms), " %.4f\n", kv.second.ms);
      os << ms;
    }
  }
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return fail(FAP_E_IO, std::string("cannot write ") + tmp);
    f << os.str();
    if (!f) return fail(FAP_E_IO, std::string("write failed: ") + tmp);
  }
  if (rename(tmp.c_str(), path) != 0) return fail(FAP_E_IO, std::string("cannot replace ") + path);
  return FAP_OK;
}

int fap_tune_cache_load(fap_tune_cache *c, const char *path) {
  if (!c || !path) return fail(FAP_E_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(FAP_E_IO, std::string("cannot read ") + path);
  std::string line;
  int lineno = 0;
  std::map<std::pair<std::string, int64_t>, fap_tune_cache::Entry> got;
  while (std::getline(f, line)) {
    lineno++;
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    std::string kind;
    if (!(is >> kind)) continue;
    const int np = nparams_of(kind);
    int64_t key;
    fap_tune_cache::Entry e;
    bool ok = np > 0 && (bool)(is >> key) && key >= 0;
    for (int i = 0; ok && i < np; i++) ok = (bool)(is >> e.p[i]);
    ok = ok && (bool)(is >> e.ms) && e.ms >= 0 && valid(kind, e.p);
    std::string extra;
    if (!ok || (is >> extra))
      return fail(FAP_E_IO, std::string(path) + ":" + std::to_string(lineno) + ": bad entry");
    got[{kind, key}] = e;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  for (auto &kv : got) c->entries[kv.first] = kv.second;
  return FAP_OK;
}

// Quarter octaves of n: the smallest k with n^4 <= 2^k (n <= 46340, so n^4
// fits in 64 bits).  The best block size changes between 5,120 and 6,144,
// one octave bucket would hold both.  A dense closure relaxes every cell in
// every round, so its cost does not depend on the matrix's density, which is
// therefore not part of the key.
int64_t fap_tune_key_fw(int32_t n) {
  const uint64_t v = (uint64_t)(n > 1 ? n : 1);
  return ceil_log2(v * v * v * v);
}

int64_t fap_tune_key_sssp(int32_t V, int64_t E, int32_t n_src) {
  const int64_t vb = ceil_log2((uint64_t)(V > 1 ? V : 1));
  const double deg = V > 0 ? (double)E / (double)V : 0.0;
  const int64_t db = (int64_t)lround(2.0 * log2(1.0 + deg));
  const int64_t sb = ceil_log2((uint64_t)(n_src > 1 ? n_src : 1));
  return ((vb * 64 + (db < 63 ? db : 63)) * 32) + (sb < 31 ? sb : 31);
}

int64_t fap_tune_key_batch(int32_t count, const int32_t *n) {
  uint64_t tiles = 0;
  for (int32_t j = 0; n && j < count; j++) {
    const uint64_t nb = (uint64_t)((n[j] > 0 ? n[j] : 0) + 63) / 64;
    tiles += nb * nb * nb;
  }
  return ceil_log2(tiles > 1 ? tiles : 1);
}

int fap_ctx_set_tune_cache(fap_ctx *ctx, fap_tune_cache *c) {
  if (!ctx) return fail(FAP_E_ARG, "ctx is null");
  ctx->tune = c;
  return FAP_OK;
}

}  // extern "C"
