// fap.cpp -- fap::fapGraph (fap/fap.h:22-125, fap/fap.cpp:27-277) on MI355X.
//
// Host side: file/CSR input, partition, preCondition (all on the CPU, as in
// the reference).  Solves: the CSR, the partition and the result blocks live
// in HBM; each part runs fap_subgraph_solve (SSSP -> gather -> FW -> min-plus
// -> scatter) on the device; getters copy blocks to the host on demand.
#include "fap/fap.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>

#include "fap_dev.h"
#include "host.h"

namespace fap {

namespace {
constexpr float kRefInf = 1e8f;  // the reference's MAXVALUE (fap/utils/parameter.h:26)

inline float to_ref_float(int32_t v) { return v >= FAP_INF ? kRefInf : (float)v; }

// Parts per batched group in solve()/solveSubGraphs(): groups alternate
// between two streams so one group's SSSP overlaps the previous group's
// closure and products.
constexpr int32_t kGroupParts = 32;
}  // namespace

struct fapGraph::Device {
  fap_ctx *ctx = nullptr;
  fap_graph_dev *g = nullptr;
  bool partition_uploaded = false;
  int32_t *d_block = nullptr;  // current part block (n_sub x V)
  size_t block_bytes = 0;
  int32_t *d_all = nullptr;    // full V x V (solve)
  size_t all_bytes = 0;
  int32_t block_id = 0;        // part held in d_block (0 = none)
  bool block_host_valid = false;
  std::vector<int32_t> host_block;  // int32 copy of d_block
  bool all_host_valid = false;
  std::vector<int32_t> host_all;
  // solveSubGraphs: one buffer, one block per part id
  int32_t *d_multi = nullptr;
  size_t multi_bytes = 0;
  std::unordered_map<int32_t, int64_t> multi_off;  // part id -> row offset in d_multi
  // predecessor rows beside each distance buffer (is_path_needed)
  int32_t *d_block_path = nullptr, *d_all_path = nullptr, *d_multi_path = nullptr;
  size_t block_path_bytes = 0, all_path_bytes = 0, multi_path_bytes = 0;
  bool block_path_valid = false, all_path_valid = false, multi_path_valid = false;

  // grows a device buffer to `need` bytes (contents not kept)
  int grow(int32_t *&p, size_t &have, size_t need) {
    if (have >= need) return FAP_OK;
    if (p) fap_dev_free(ctx, p);
    p = nullptr;
    have = 0;
    void *q = nullptr;
    const int rc = fap_dev_alloc(ctx, need, &q);
    if (rc != FAP_OK) return rc;
    p = (int32_t *)q;
    have = need;
    return FAP_OK;
  }

  ~Device() {
    if (ctx) {
      int32_t *bufs[] = {d_block, d_all, d_multi, d_block_path, d_all_path, d_multi_path};
      for (int32_t *b : bufs)
        if (b) fap_dev_free(ctx, b);
      if (g) fap_graph_free(ctx, g);
      fap_ctx_destroy(ctx);
    }
  }
};

void fapGraph::reset_status() {
  status_ = 0;
  error_.clear();
}

int32_t fapGraph::fail(int32_t code, const std::string &msg) {
  status_ = code;
  error_ = msg;
  return -1;
}

void fapGraph::init_from_csr(int32_t V, std::vector<int32_t> &&ro, std::vector<int32_t> &&col,
                             std::vector<int32_t> &&w, bool directed_graph) {
  num_vertexs = V;
  num_edges = (int32_t)col.size();
  directed = directed_graph;
  row_offset = std::move(ro);
  col_val = std::move(col);
  weight_i32_ = std::move(w);
  adj_size.resize(V);
  for (int32_t v = 0; v < V; v++) adj_size[v] = row_offset[v + 1] - row_offset[v];
  weight.resize(weight_i32_.size());
  for (size_t e = 0; e < weight_i32_.size(); e++) weight[e] = (float)weight_i32_[e];
  graph_id.assign(V, 0);
  st2ed.assign(V, 0);
  ed2st.assign(V, 0);
  C_BlockVer_num.assign(K + 1, 0);
  C_BlockVer_offset.assign(K + 1, 0);
  C_BlockBdy_num.assign(K + 1, 0);
  C_BlockBdy_offset.assign(K + 1, 0);
}

fapGraph::fapGraph(std::string input_graph_, bool directed_, bool weighted_, int32_t K_) {
  input_graph = input_graph_;
  directed = directed_;
  weighted = weighted_;
  K = K_;
  fapi::HostCSR csr;
  std::string err;
  const std::string path = fapi::reference_graph_path(input_graph_, directed_, weighted_);
  int rc = fapi::read_mtx(path, directed_, weighted_, &csr, &err);
  if (rc != FAP_OK) {
    fail(rc, err);
    return;
  }
  if (K_ < 1 || K_ > std::max<int32_t>(csr.V, 1)) {
    fail(FAP_E_ARG, "K must be in 1..V");
    return;
  }
  std::vector<int32_t> gid;
  rc = fapi::partition_graph(csr, directed_, K_, &gid, &used_metis_, &err);
  init_from_csr(csr.V, std::move(csr.row_offset), std::move(csr.col), std::move(csr.w),
                directed_);
  if (rc != FAP_OK) {
    fail(rc, err);
    return;
  }
  graph_id = gid;
  for (int32_t i = 0; i < num_vertexs; i++) BlockVer[graph_id[i]].push_back(i);
  loaded_ = true;
}

fapGraph::fapGraph(int32_t V, int64_t E, std::vector<int32_t> ro, std::vector<int32_t> col,
                   std::vector<float> w, int32_t K_) {
  K = K_;
  weighted = true;
  if (V < 0 || E < 0 || (int64_t)ro.size() != (int64_t)V + 1 || (int64_t)col.size() != E ||
      (int64_t)w.size() != E || E > INT32_MAX) {
    fail(FAP_E_ARG, "CSR sizes inconsistent");
    return;
  }
  if (ro[0] != 0 || ro[V] != E) {
    fail(FAP_E_ARG, "row_offset must run from 0 to num_edges");
    return;
  }
  std::vector<int32_t> wi(E);
  for (int64_t e = 0; e < E; e++) {
    const float x = w[e];
    if (!(x >= 0.f) || x != floorf(x) || x >= (float)FAP_INF) {
      fail(FAP_E_ARG, "weights must be non-negative integers (int32 build)");
      return;
    }
    if (col[e] < 0 || col[e] >= V) {
      fail(FAP_E_ARG, "col_val entry out of range");
      return;
    }
    wi[e] = (int32_t)x;
  }
  for (int32_t v = 0; v < V; v++)
    if (ro[v + 1] < ro[v]) {
      fail(FAP_E_ARG, "row_offset not monotone");
      return;
    }
  // undirected iff every arc u->v has a reverse arc v->u
  bool sym = true;
  {
    std::vector<std::pair<int32_t, int32_t>> a, b;
    a.reserve(E);
    b.reserve(E);
    for (int32_t u = 0; u < V; u++)
      for (int32_t e = ro[u]; e < ro[u + 1]; e++) {
        a.emplace_back(u, col[e]);
        b.emplace_back(col[e], u);
      }
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    sym = (a == b);
  }
  if (K_ < 1 || K_ > std::max<int32_t>(V, 1)) {
    init_from_csr(V, std::move(ro), std::move(col), std::move(wi), !sym);
    fail(FAP_E_ARG, "K must be in 1..V");
    return;
  }
  fapi::HostCSR csr;
  csr.V = V;
  csr.E = E;
  csr.row_offset = ro;
  csr.col = col;
  csr.w = wi;
  std::vector<int32_t> gid;
  std::string err;
  int rc = fapi::partition_graph(csr, !sym, K_, &gid, &used_metis_, &err);
  init_from_csr(V, std::move(ro), std::move(col), std::move(wi), !sym);
  if (rc != FAP_OK) {
    fail(rc, err);
    return;
  }
  graph_id = gid;
  for (int32_t i = 0; i < num_vertexs; i++) BlockVer[graph_id[i]].push_back(i);
  loaded_ = true;
}

fapGraph::~fapGraph() = default;

int32_t fapGraph::setPartition(const std::vector<int32_t> &gid) {
  if (!loaded_ && num_vertexs == 0 && row_offset.empty()) return fail(FAP_E_STATE, "graph not loaded");
  if ((int32_t)gid.size() != num_vertexs) return fail(FAP_E_ARG, "partition size != V");
  for (int32_t v : gid)
    if (v < 1 || v > K) return fail(FAP_E_ARG, "part id outside 1..K");
  graph_id = gid;
  BlockVer.clear();
  for (int32_t i = 0; i < num_vertexs; i++) BlockVer[graph_id[i]].push_back(i);
  preconditioned_ = false;
  full_solved_ = false;
  if (dev_) {
    dev_->partition_uploaded = false;
    dev_->block_id = 0;
    dev_->multi_off.clear();
  }
  used_metis_ = false;
  loaded_ = true;
  status_ = 0;
  error_.clear();
  return 0;
}

// fap/fap.cpp:67-81
void fapGraph::preCondition() {
  reset_status();
  if (!loaded_) {
    fail(FAP_E_STATE, "graph not loaded: " + error_);
    return;
  }
  fapi::HostCSR csr;
  csr.V = num_vertexs;
  csr.E = num_edges;
  csr.row_offset = row_offset;
  csr.col = col_val;
  fapi::PartInfo pi;
  fapi::precondition(csr, K, graph_id, &pi);
  st2ed = pi.st2ed;
  ed2st = pi.ed2st;
  C_BlockVer_num = pi.ver_num;
  C_BlockVer_offset = pi.ver_off;
  C_BlockBdy_num = pi.bdy_num;
  C_BlockBdy_offset = pi.bdy_off;
  isBoundry.clear();
  BlockBoundary.clear();
  for (int32_t i = 0; i < num_vertexs; i++) {
    if (pi.is_bdy[i]) {
      isBoundry[i] = 1;
      BlockBoundary[graph_id[i]].push_back(i);
    }
  }
  max_subgraph_vertexs = 0;
  for (int32_t p = 1; p <= K; p++) max_subgraph_vertexs = std::max(max_subgraph_vertexs, C_BlockVer_num[p]);
  is_split = (num_vertexs > max_vertexs_num_limit);
  preconditioned_ = true;
  full_solved_ = false;
  if (dev_) {
    dev_->partition_uploaded = false;
    dev_->block_id = 0;
    dev_->multi_off.clear();
  }
}

bool fapGraph::isSplit() { return is_split; }

std::vector<int> fapGraph::getGraphId() { return graph_id; }

std::vector<int32_t> fapGraph::getBoundaryFlags() const {
  std::vector<int32_t> f(num_vertexs, 0);
  for (auto &kv : isBoundry)
    if (kv.first >= 0 && kv.first < num_vertexs) f[kv.first] = kv.second;
  return f;
}

int32_t fapGraph::ensure_device() {
  if (!dev_) dev_.reset(new Device());
  Device &d = *dev_;
  if (!d.ctx) {
    if (fap_ctx_create(0, &d.ctx) != FAP_OK) {
      d.ctx = nullptr;
      return fail(FAP_E_HIP, std::string("GPU context: ") + fap_last_error());
    }
  }
  if (!d.g) {
    if (fap_graph_upload(d.ctx, num_vertexs, num_edges, row_offset.data(), col_val.data(),
                         weight_i32_.data(), &d.g) != FAP_OK) {
      d.g = nullptr;
      return fail(FAP_E_HIP, std::string("graph upload: ") + fap_last_error());
    }
  }
  if (!d.partition_uploaded) {
    if (fap_graph_set_partition(d.ctx, d.g, graph_id.data(), st2ed.data(), ed2st.data()) !=
        FAP_OK)
      return fail(FAP_E_HIP, std::string("partition upload: ") + fap_last_error());
    d.partition_uploaded = true;
  }
  return 0;
}

bool fapGraph::part_ok(int32_t id) const {
  return preconditioned_ && id >= 1 && id <= K;
}

// fap/fap.cpp:92-181, device resident.
int32_t fapGraph::run(int32_t id, int32_t *d_rowblock) {
  const int32_t n_sub = C_BlockVer_num[id];
  if (n_sub == 0) return 0;
  if (fap_subgraph_solve(dev_->ctx, dev_->g, C_BlockVer_offset[id], n_sub, C_BlockBdy_num[id],
                         d_rowblock, num_vertexs, nullptr) != FAP_OK)
    return fail(FAP_E_HIP, std::string("part ") + std::to_string(id) + ": " + fap_last_error());
  return 0;
}

// Batched form of run() over several parts (fap_subgraph_solve_batch).
int32_t fapGraph::run_batch(const std::vector<int32_t> &ids, const std::vector<int32_t *> &blocks) {
  std::vector<int32_t> starts, ns, bs;
  std::vector<int32_t *> rb;
  for (size_t i = 0; i < ids.size(); i++) {
    const int32_t id = ids[i];
    if (C_BlockVer_num[id] == 0) continue;
    starts.push_back(C_BlockVer_offset[id]);
    ns.push_back(C_BlockVer_num[id]);
    bs.push_back(C_BlockBdy_num[id]);
    rb.push_back(blocks[i]);
  }
  if (starts.empty()) return 0;
  if (fap_subgraph_solve_groups(dev_->ctx, dev_->g, (int32_t)starts.size(), starts.data(),
                                ns.data(), bs.data(), rb.data(), num_vertexs, kGroupParts) != FAP_OK)
    return fail(FAP_E_HIP, std::string("batched parts: ") + fap_last_error());
  return 0;
}

// Predecessor rows of one solved part block (fap_subgraph_pred).
int32_t fapGraph::run_pred(int32_t id, const int32_t *d_rowblock, int32_t *d_predblock) {
  if (C_BlockVer_num[id] == 0) return 0;
  if (fap_subgraph_pred(dev_->ctx, dev_->g, C_BlockVer_offset[id], C_BlockVer_num[id], d_rowblock,
                        num_vertexs, d_predblock, num_vertexs, nullptr) != FAP_OK)
    return fail(FAP_E_HIP, std::string("paths of part ") + std::to_string(id) + ": " +
                               fap_last_error());
  return 0;
}

// fap/fap.cpp:184-194.  The full V x V matrix, parts stacked in part order.
int32_t fapGraph::solve(bool is_path_needed_) {
  reset_status();
  is_path_needed = is_path_needed_;
  if (!preconditioned_) return fail(FAP_E_STATE, "solve() before preCondition()");
  if (ensure_device() != 0) return -1;
  Device &d = *dev_;
  const size_t bytes = (size_t)num_vertexs * num_vertexs * sizeof(int32_t);
  if (d.all_bytes < bytes) {
    if (d.d_all) fap_dev_free(d.ctx, d.d_all);
    d.d_all = nullptr;
    d.all_bytes = 0;
    void *p = nullptr;
    if (fap_dev_alloc(d.ctx, bytes, &p) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("the V x V matrix does not fit: ") + fap_last_error());
    d.d_all = (int32_t *)p;
    d.all_bytes = bytes;
  }
  if (batched_) {
    std::vector<int32_t> ids;
    std::vector<int32_t *> blocks;
    for (int32_t i = 1; i <= K; i++) {
      ids.push_back(i);
      blocks.push_back(d.d_all + (int64_t)C_BlockVer_offset[i] * num_vertexs);
    }
    if (run_batch(ids, blocks) != 0) return -1;
  } else {
    for (int32_t i = 1; i <= K; i++) {
      if (run(i, d.d_all + (int64_t)C_BlockVer_offset[i] * num_vertexs) != 0) return -1;
    }
  }
  if (fap_ctx_sync(d.ctx) != FAP_OK) return fail(FAP_E_HIP, fap_last_error());
  d.all_path_valid = false;
  if (is_path_needed) {
    if (d.grow(d.d_all_path, d.all_path_bytes, bytes) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("the V x V path matrix does not fit: ") +
                                   fap_last_error());
    for (int32_t i = 1; i <= K; i++) {
      const int64_t off = (int64_t)C_BlockVer_offset[i] * num_vertexs;
      if (run_pred(i, d.d_all + off, d.d_all_path + off) != 0) return -1;
    }
    d.all_path_valid = true;
  }
  full_solved_ = true;
  d.all_host_valid = false;
  current_subgraph_id = K;
  status_ = 0;
  error_.clear();
  return 0;
}

// fap/fap.cpp:197-207
int32_t fapGraph::solveSubGraph(int32_t id, bool is_path_needed_) {
  reset_status();
  is_path_needed = is_path_needed_;
  if (!part_ok(id)) return fail(FAP_E_ARG, "part id outside 1..K or preCondition() not run");
  if (ensure_device() != 0) return -1;
  Device &d = *dev_;
  const size_t bytes = (size_t)C_BlockVer_num[id] * num_vertexs * sizeof(int32_t);
  if (d.block_bytes < bytes) {
    if (d.d_block) fap_dev_free(d.ctx, d.d_block);
    d.d_block = nullptr;
    d.block_bytes = 0;
    void *p = nullptr;
    if (fap_dev_alloc(d.ctx, bytes, &p) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("part block: ") + fap_last_error());
    d.d_block = (int32_t *)p;
    d.block_bytes = bytes;
  }
  d.block_id = 0;
  d.block_path_valid = false;
  if ((batched_ ? run_batch({id}, {d.d_block}) : run(id, d.d_block)) != 0) return -1;
  if (fap_ctx_sync(d.ctx) != FAP_OK) return fail(FAP_E_HIP, fap_last_error());
  if (is_path_needed) {
    if (d.grow(d.d_block_path, d.block_path_bytes, bytes) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("part path block: ") + fap_last_error());
    if (run_pred(id, d.d_block, d.d_block_path) != 0) return -1;
    d.block_path_valid = true;
  }
  d.block_id = id;
  d.block_host_valid = false;
  current_subgraph_id = id;
  status_ = 0;
  error_.clear();
  return 0;
}

int32_t fapGraph::solveSubGraphs(const std::vector<int32_t> &ids, bool is_path_needed_) {
  reset_status();
  is_path_needed = is_path_needed_;
  for (int32_t id : ids)
    if (!part_ok(id)) return fail(FAP_E_ARG, "part id outside 1..K or preCondition() not run");
  if (ensure_device() != 0) return -1;
  Device &d = *dev_;
  int64_t rows = 0;
  std::unordered_map<int32_t, int64_t> off;
  for (int32_t id : ids) {
    if (off.count(id)) continue;
    off[id] = rows;
    rows += C_BlockVer_num[id];
  }
  const size_t bytes = (size_t)std::max<int64_t>(rows, 1) * num_vertexs * sizeof(int32_t);
  if (d.multi_bytes < bytes) {
    if (d.d_multi) fap_dev_free(d.ctx, d.d_multi);
    d.d_multi = nullptr;
    d.multi_bytes = 0;
    void *p = nullptr;
    if (fap_dev_alloc(d.ctx, bytes, &p) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("part blocks: ") + fap_last_error());
    d.d_multi = (int32_t *)p;
    d.multi_bytes = bytes;
  }
  d.multi_off.clear();
  std::vector<int32_t> uids;
  std::vector<int32_t *> blocks;
  for (auto &kv : off) {
    uids.push_back(kv.first);
    blocks.push_back(d.d_multi + kv.second * num_vertexs);
  }
  if (batched_) {
    if (run_batch(uids, blocks) != 0) return -1;
  } else {
    for (size_t i = 0; i < uids.size(); i++)
      if (run(uids[i], blocks[i]) != 0) return -1;
    if (fap_ctx_sync(d.ctx) != FAP_OK) return fail(FAP_E_HIP, fap_last_error());
  }
  d.multi_path_valid = false;
  if (is_path_needed) {
    if (d.grow(d.d_multi_path, d.multi_path_bytes, bytes) != FAP_OK)
      return fail(FAP_E_NOMEM, std::string("part path blocks: ") + fap_last_error());
    for (size_t i = 0; i < uids.size(); i++)
      if (run_pred(uids[i], blocks[i], d.d_multi_path + off[uids[i]] * num_vertexs) != 0)
        return -1;
    d.multi_path_valid = true;
  }
  d.multi_off = off;
  current_subgraph_id = ids.empty() ? current_subgraph_id : ids.back();
  return 0;
}

std::vector<int32_t> fapGraph::getMapping() { return st2ed; }

std::vector<int32_t> fapGraph::getAllDistanceInt32() {
  reset_status();
  if (!full_solved_ || !dev_) {
    fail(FAP_E_STATE, "solve() has not produced the full matrix");
    return {};
  }
  Device &d = *dev_;
  if (!d.all_host_valid) {
    d.host_all.resize((size_t)num_vertexs * num_vertexs);
    if (fap_memcpy_d2h(d.ctx, d.host_all.data(), d.d_all, d.host_all.size() * sizeof(int32_t)) !=
        FAP_OK) {
      fail(FAP_E_HIP, fap_last_error());
      return {};
    }
    d.all_host_valid = true;
  }
  return d.host_all;
}

std::vector<float> fapGraph::getAllDistance() {
  std::vector<int32_t> v = getAllDistanceInt32();
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); i++) f[i] = to_ref_float(v[i]);
  return f;
}

std::vector<int32_t> fapGraph::getAllPath() {
  reset_status();
  if (!full_solved_ || !dev_ || !dev_->all_path_valid) {
    fail(FAP_E_STATE, "solve(true) has not produced the full path matrix");
    return {};
  }
  path.resize((size_t)num_vertexs * num_vertexs);
  if (!path.empty() && fap_memcpy_d2h(dev_->ctx, path.data(), dev_->d_all_path,
                                      path.size() * sizeof(int32_t)) != FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  return path;
}

const int32_t *fapGraph::subGraphPathDevice(int32_t id) {
  if (!dev_ || !part_ok(id)) return nullptr;
  Device &d = *dev_;
  if (d.block_id == id) return d.block_path_valid ? d.d_block_path : nullptr;
  auto it = d.multi_off.find(id);
  if (it != d.multi_off.end()) return d.multi_path_valid ? d.d_multi_path + it->second * num_vertexs : nullptr;
  if (full_solved_ && d.all_path_valid)
    return d.d_all_path + (int64_t)C_BlockVer_offset[id] * num_vertexs;
  return nullptr;
}

int32_t fapGraph::getCurrentSubGraphId() { return current_subgraph_id; }

std::vector<int32_t> fapGraph::getSubGraphIndex(int32_t id) {
  if (!part_ok(id)) return {};
  const int off = C_BlockVer_offset[id];
  const int n = C_BlockVer_num[id];
  return std::vector<int32_t>(st2ed.begin() + off, st2ed.begin() + off + n);
}

const int32_t *fapGraph::subGraphDistanceDevice(int32_t id, int64_t *ld) {
  if (ld) *ld = num_vertexs;
  if (!dev_ || !part_ok(id)) return nullptr;
  if (dev_->block_id == id) return dev_->d_block;
  auto it = dev_->multi_off.find(id);
  if (it != dev_->multi_off.end()) return dev_->d_multi + it->second * num_vertexs;
  if (full_solved_) return dev_->d_all + (int64_t)C_BlockVer_offset[id] * num_vertexs;
  return nullptr;
}

std::vector<int32_t> fapGraph::getSubGraphDistanceInt32(int32_t id) {
  reset_status();
  int64_t ld = 0;
  const int32_t *p = subGraphDistanceDevice(id, &ld);
  if (!p) {
    fail(FAP_E_STATE, "part " + std::to_string(id) + " has not been solved");
    return {};
  }
  Device &d = *dev_;
  const size_t n = (size_t)C_BlockVer_num[id] * num_vertexs;
  if (p == d.d_block && d.block_host_valid) return d.host_block;
  std::vector<int32_t> out(n);
  if (n && fap_memcpy_d2h(d.ctx, out.data(), p, n * sizeof(int32_t)) != FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  if (p == d.d_block) {
    d.host_block = out;
    d.block_host_valid = true;
  }
  return out;
}

std::vector<float> fapGraph::getSubGraphDistance(int32_t id) {
  std::vector<int32_t> v = getSubGraphDistanceInt32(id);
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); i++) f[i] = to_ref_float(v[i]);
  subgraph_dist = f;
  return f;
}

std::vector<int32_t> fapGraph::getSubGraphPath(int32_t id) {
  reset_status();
  const int32_t *p = subGraphPathDevice(id);
  if (!p) {
    fail(FAP_E_STATE, "part " + std::to_string(id) + " has no path rows (solve with is_path_needed)");
    return {};
  }
  std::vector<int32_t> out((size_t)C_BlockVer_num[id] * num_vertexs);
  if (!out.empty() &&
      fap_memcpy_d2h(dev_->ctx, out.data(), p, out.size() * sizeof(int32_t)) != FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  subgraph_path = out;
  return out;
}

int32_t fapGraph::getSubGraphRowsInt32(int32_t id, int32_t row0, int32_t nrows, int32_t *out) {
  reset_status();
  if (!part_ok(id) || row0 < 0 || nrows < 0 || (int64_t)row0 + nrows > C_BlockVer_num[id] ||
      (nrows > 0 && !out))
    return fail(FAP_E_ARG, "row range outside the part");
  int64_t ld = 0;
  const int32_t *p = subGraphDistanceDevice(id, &ld);
  if (!p) return fail(FAP_E_STATE, "part " + std::to_string(id) + " has not been solved");
  for (int32_t r = 0; r < nrows; r++)
    if (fap_memcpy_d2h(dev_->ctx, out + (int64_t)r * num_vertexs, p + (row0 + r) * ld,
                       sizeof(int32_t) * num_vertexs) != FAP_OK)
      return fail(FAP_E_HIP, fap_last_error());
  status_ = 0;
  error_.clear();
  return 0;
}

std::vector<float> fapGraph::getDistanceFromOnePoint(int32_t v) {
  reset_status();
  if (v < 0 || v >= num_vertexs || !part_ok(current_subgraph_id) ||
      graph_id[v] != current_subgraph_id) {
    fail(FAP_E_ARG, "vertex is not in the current part");
    return {};
  }
  int64_t ld = 0;
  const int32_t *p = subGraphDistanceDevice(current_subgraph_id, &ld);
  if (!p) {
    fail(FAP_E_STATE, "current part not solved");
    return {};
  }
  const int64_t row = ed2st[v] - C_BlockVer_offset[current_subgraph_id];
  std::vector<int32_t> tmp(num_vertexs);
  if (fap_memcpy_d2h(dev_->ctx, tmp.data(), p + row * ld, tmp.size() * sizeof(int32_t)) !=
      FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  std::vector<float> f(num_vertexs);
  for (int32_t i = 0; i < num_vertexs; i++) f[i] = to_ref_float(tmp[i]);
  return f;
}

std::vector<int32_t> fapGraph::getPathFromOnePoint(int32_t v) {
  reset_status();
  if (v < 0 || v >= num_vertexs || !part_ok(current_subgraph_id) ||
      graph_id[v] != current_subgraph_id) {
    fail(FAP_E_ARG, "vertex is not in the current part");
    return {};
  }
  const int32_t *p = subGraphPathDevice(current_subgraph_id);
  if (!p) {
    fail(FAP_E_STATE, "current part has no path rows (solve with is_path_needed)");
    return {};
  }
  const int64_t row = ed2st[v] - C_BlockVer_offset[current_subgraph_id];
  std::vector<int32_t> out(num_vertexs);
  if (fap_memcpy_d2h(dev_->ctx, out.data(), p + row * num_vertexs,
                     out.size() * sizeof(int32_t)) != FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  return out;
}

float fapGraph::getDistanceP2P() {
  fail(FAP_E_ARG, "getDistanceP2P() needs (src, dst)");
  return -1.f;
}

int32_t fapGraph::getDistanceP2P(int32_t src, int32_t dst, int32_t *out) {
  reset_status();
  if (!out || src < 0 || src >= num_vertexs || dst < 0 || dst >= num_vertexs)
    return fail(FAP_E_ARG, "bad vertex id");
  if (!preconditioned_) return fail(FAP_E_STATE, "preCondition() not run");
  const int32_t part = graph_id[src];
  int64_t ld = 0;
  const int32_t *p = subGraphDistanceDevice(part, &ld);
  if (!p) return fail(FAP_E_STATE, "the part of src has not been solved");
  const int64_t row = ed2st[src] - C_BlockVer_offset[part];
  if (fap_memcpy_d2h(dev_->ctx, out, p + row * ld + dst, sizeof(int32_t)) != FAP_OK)
    return fail(FAP_E_HIP, fap_last_error());
  return 0;
}

std::vector<int32_t> fapGraph::getPathP2P() {
  fail(FAP_E_ARG, "getPathP2P() needs (src, dst)");
  return {};
}

std::vector<int32_t> fapGraph::getPathP2P(int32_t src, int32_t dst) {
  reset_status();
  if (src < 0 || src >= num_vertexs || dst < 0 || dst >= num_vertexs) {
    fail(FAP_E_ARG, "bad vertex id");
    return {};
  }
  if (!preconditioned_) {
    fail(FAP_E_STATE, "preCondition() not run");
    return {};
  }
  const int32_t part = graph_id[src];
  const int32_t *p = subGraphPathDevice(part);
  if (!p) {
    fail(FAP_E_STATE, "the part of src has no path rows (solve with is_path_needed)");
    return {};
  }
  const int64_t row = ed2st[src] - C_BlockVer_offset[part];
  std::vector<int32_t> pr(num_vertexs);
  if (fap_memcpy_d2h(dev_->ctx, pr.data(), p + row * num_vertexs, pr.size() * sizeof(int32_t)) !=
      FAP_OK) {
    fail(FAP_E_HIP, fap_last_error());
    return {};
  }
  if (pr[dst] < 0) return {};
  std::vector<int32_t> seq{dst};
  for (int32_t x = dst, steps = 0; x != src; steps++) {
    x = pr[x];
    if (x < 0 || x >= num_vertexs || steps >= num_vertexs) {
      fail(FAP_E_STATE, "path rows are inconsistent");
      return {};
    }
    seq.push_back(x);
  }
  std::reverse(seq.begin(), seq.end());
  return seq;
}

int32_t fapGraph::subGraphChecksum(int32_t id, uint64_t *sum, uint64_t *weighted_sum) {
  reset_status();
  int64_t ld = 0;
  const int32_t *p = subGraphDistanceDevice(id, &ld);
  if (!p) return fail(FAP_E_STATE, "part not solved");
  if (fap_checksum_i32(dev_->ctx, p, (int64_t)C_BlockVer_num[id] * num_vertexs, sum, weighted_sum,
                       nullptr) != FAP_OK)
    return fail(FAP_E_HIP, fap_last_error());
  return 0;
}

// Restates check_ans (fap/utils/checkResult.h:38-82) as a deterministic
// verification helper: 20 sampled rows x 10 sampled columns of the given
// block against a host Dijkstra from the row's vertex, pairs with both
// values above MAXVALUE/10 skipped, tolerance 1e-5.  The samples come from a
// fixed splitmix64 sequence instead of time(NULL); the path argument is not
// read (the reference reads path entries that may be -1, :57,62).
bool fapGraph::check_result(float *dist_, int *, int *source, int source_num, int vertexs,
                            int *adj_size_, int *row_offset_, int *col_val_, float *weight_,
                            int *) {
  if (!dist_ || !source || source_num <= 0 || vertexs <= 0) return false;
  uint64_t state = 0x243F6A8885A308D3ull;
  auto next = [&]() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  std::vector<float> ref(vertexs);
  for (int c = 0; c < 20; c++) {
    const int idx = (int)(next() % (uint64_t)source_num);
    const int src = source[idx];
    std::fill(ref.begin(), ref.end(), kRefInf);
    std::vector<char> done(vertexs, 0);
    typedef std::pair<float, int> Item;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    ref[src] = 0;
    heap.push(Item(0.f, src));
    while (!heap.empty()) {
      Item it = heap.top();
      heap.pop();
      if (done[it.second]) continue;
      done[it.second] = 1;
      const int u = it.second;
      for (int i = 0; i < adj_size_[u]; i++) {
        const int v = col_val_[row_offset_[u] + i];
        const float nd = it.first + weight_[row_offset_[u] + i];
        if (ref[v] > nd) {
          ref[v] = nd;
          heap.push(Item(nd, v));
        }
      }
    }
    for (int k = 0; k < 10; k++) {
      const int dst = (int)(next() % (uint64_t)vertexs);
      const float got = dist_[(int64_t)idx * vertexs + dst];
      if (got > kRefInf / 10 && ref[dst] > kRefInf / 10) continue;
      if (fabsf(got - ref[dst]) > 1e-5f) return false;
    }
  }
  return true;
}

}  // namespace fap
