// capi_graph.cpp -- C ABI over fap::fapGraph (include/fap_graph.h).
#include <string.h>

#include <string>
#include <vector>

#include "common.h"
#include "fap/fap.h"
#include "fap_graph.h"

struct fapg {
  fap::fapGraph *g = nullptr;
};

using fapi::fail;

namespace {

int from_status(fap::fapGraph *g) {
  if (g->status() == 0) return FAP_OK;
  return fail(g->status() < 0 ? g->status() : FAP_E_STATE, g->lastError());
}

template <typename T>
int copy_out(const std::vector<T> &v, T *out, const char *what) {
  if (!out) return fail(FAP_E_ARG, std::string(what) + ": null output");
  if (!v.empty()) memcpy(out, v.data(), v.size() * sizeof(T));
  return FAP_OK;
}

}  // namespace

extern "C" {

int fapg_create_from_file(const char *name, int directed, int weighted, int32_t K, fapg **out) {
  if (!name || !out) return fail(FAP_E_ARG, "null argument");
  fapg *h = new fapg();
  h->g = new fap::fapGraph(std::string(name), directed != 0, weighted != 0, K);
  if (h->g->status() != 0) {
    int rc = from_status(h->g);
    delete h->g;
    delete h;
    return rc;
  }
  *out = h;
  fapi::clear_error();
  return FAP_OK;
}

int fapg_create_from_csr(int32_t V, int64_t E, const int32_t *row_offset, const int32_t *col,
                         const float *weight, int32_t K, fapg **out) {
  if (!out || !row_offset || V < 0 || E < 0 || (E > 0 && (!col || !weight)))
    return fail(FAP_E_ARG, "null argument");
  std::vector<int32_t> ro(row_offset, row_offset + V + 1);
  std::vector<int32_t> c(col, col + E);
  std::vector<float> w(weight, weight + E);
  fapg *h = new fapg();
  h->g = new fap::fapGraph(V, E, std::move(ro), std::move(c), std::move(w), K);
  if (h->g->status() != 0) {
    int rc = from_status(h->g);
    delete h->g;
    delete h;
    return rc;
  }
  *out = h;
  fapi::clear_error();
  return FAP_OK;
}

int fapg_destroy(fapg *h) {
  if (!h) return FAP_OK;
  delete h->g;
  delete h;
  return FAP_OK;
}

#define FAPG_CHECK(h)                                         \
  do {                                                        \
    if (!(h) || !(h)->g) return fail(FAP_E_ARG, "null graph"); \
  } while (0)

int fapg_set_partition(fapg *h, const int32_t *gid) {
  FAPG_CHECK(h);
  if (!gid) return fail(FAP_E_ARG, "null graph_id");
  std::vector<int32_t> v(gid, gid + h->g->num_vertexs);
  if (h->g->setPartition(v) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_precondition(fapg *h) {
  FAPG_CHECK(h);
  h->g->preCondition();
  return from_status(h->g);
}

int fapg_is_split(fapg *h, int *out) {
  FAPG_CHECK(h);
  if (!out) return fail(FAP_E_ARG, "null output");
  *out = h->g->isSplit() ? 1 : 0;
  return FAP_OK;
}

int fapg_solve(fapg *h, int is_path_needed) {
  FAPG_CHECK(h);
  if (h->g->solve(is_path_needed != 0) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_solve_subgraph(fapg *h, int32_t id, int is_path_needed) {
  FAPG_CHECK(h);
  if (h->g->solveSubGraph(id, is_path_needed != 0) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_set_batched(fapg *h, int batched) {
  FAPG_CHECK(h);
  h->g->setBatched(batched != 0);
  return FAP_OK;
}

int fapg_solve_subgraphs(fapg *h, int32_t count, const int32_t *ids, int is_path_needed) {
  FAPG_CHECK(h);
  if (count < 0 || (count > 0 && !ids)) return fail(FAP_E_ARG, "bad id list");
  std::vector<int32_t> v(ids, ids + count);
  if (h->g->solveSubGraphs(v, is_path_needed != 0) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_num_vertices(fapg *h, int32_t *V) {
  FAPG_CHECK(h);
  if (!V) return fail(FAP_E_ARG, "null output");
  *V = h->g->num_vertexs;
  return FAP_OK;
}

int fapg_num_edges(fapg *h, int64_t *E) {
  FAPG_CHECK(h);
  if (!E) return fail(FAP_E_ARG, "null output");
  *E = h->g->num_edges;
  return FAP_OK;
}

int fapg_num_parts(fapg *h, int32_t *K) {
  FAPG_CHECK(h);
  if (!K) return fail(FAP_E_ARG, "null output");
  *K = h->g->numParts();
  return FAP_OK;
}

int fapg_is_directed(fapg *h, int *d) {
  FAPG_CHECK(h);
  if (!d) return fail(FAP_E_ARG, "null output");
  *d = h->g->directed ? 1 : 0;
  return FAP_OK;
}

int fapg_used_metis(fapg *h, int *u) {
  FAPG_CHECK(h);
  if (!u) return fail(FAP_E_ARG, "null output");
  *u = h->g->usedMetis() ? 1 : 0;
  return FAP_OK;
}

int fapg_get_csr(fapg *h, int32_t *ro, int32_t *col, int32_t *w) {
  FAPG_CHECK(h);
  FAP_TRY(copy_out(h->g->row_offset, ro, "row_offset"));
  FAP_TRY(copy_out(h->g->col_val, col, "col"));
  return copy_out(h->g->weightInt32(), w, "weight");
}

int fapg_get_graph_id(fapg *h, int32_t *out) {
  FAPG_CHECK(h);
  return copy_out(h->g->getGraphId(), out, "graph_id");
}

int fapg_get_boundary_flags(fapg *h, int32_t *out) {
  FAPG_CHECK(h);
  return copy_out(h->g->getBoundaryFlags(), out, "boundary flags");
}

int fapg_get_mapping(fapg *h, int32_t *st2ed, int32_t *ed2st) {
  FAPG_CHECK(h);
  std::vector<int32_t> m = h->g->getMapping();
  FAP_TRY(copy_out(m, st2ed, "st2ed"));
  if (ed2st)
    for (size_t i = 0; i < m.size(); i++) ed2st[m[i]] = (int32_t)i;
  return FAP_OK;
}

int fapg_get_part_info(fapg *h, int32_t *vn, int32_t *vo, int32_t *bn, int32_t *bo) {
  FAPG_CHECK(h);
  if (vn) FAP_TRY(copy_out(h->g->getPartVertexNum(), vn, "ver_num"));
  if (vo) FAP_TRY(copy_out(h->g->getPartVertexOffset(), vo, "ver_off"));
  if (bn) FAP_TRY(copy_out(h->g->getPartBoundaryNum(), bn, "bdy_num"));
  if (bo) FAP_TRY(copy_out(h->g->getPartVertexOffset(), bo, "bdy_off"));
  return FAP_OK;
}

int fapg_get_current_subgraph_id(fapg *h, int32_t *id) {
  FAPG_CHECK(h);
  if (!id) return fail(FAP_E_ARG, "null output");
  *id = h->g->getCurrentSubGraphId();
  return FAP_OK;
}

int fapg_get_subgraph_index(fapg *h, int32_t id, int32_t *out) {
  FAPG_CHECK(h);
  if (id < 1 || id > h->g->numParts()) return fail(FAP_E_ARG, "part id outside 1..K");
  return copy_out(h->g->getSubGraphIndex(id), out, "subgraph index");
}

int fapg_get_subgraph_distance_i32(fapg *h, int32_t id, int32_t *out) {
  FAPG_CHECK(h);
  std::vector<int32_t> v = h->g->getSubGraphDistanceInt32(id);
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "distance");
}

int fapg_get_subgraph_rows_i32(fapg *h, int32_t id, int32_t row0, int32_t nrows, int32_t *out) {
  FAPG_CHECK(h);
  if (h->g->getSubGraphRowsInt32(id, row0, nrows, out) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_get_subgraph_distance_f32(fapg *h, int32_t id, float *out) {
  FAPG_CHECK(h);
  std::vector<float> v = h->g->getSubGraphDistance(id);
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "distance");
}

int fapg_get_all_distance_i32(fapg *h, int32_t *out) {
  FAPG_CHECK(h);
  std::vector<int32_t> v = h->g->getAllDistanceInt32();
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "distance");
}

int fapg_get_all_distance_f32(fapg *h, float *out) {
  FAPG_CHECK(h);
  std::vector<float> v = h->g->getAllDistance();
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "distance");
}

int fapg_get_distance_from_one_point(fapg *h, int32_t v, float *out) {
  FAPG_CHECK(h);
  std::vector<float> r = h->g->getDistanceFromOnePoint(v);
  if (r.empty()) return from_status(h->g);
  return copy_out(r, out, "row");
}

int fapg_distance_p2p(fapg *h, int32_t src, int32_t dst, int32_t *out) {
  FAPG_CHECK(h);
  if (h->g->getDistanceP2P(src, dst, out) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_get_subgraph_path(fapg *h, int32_t id, int32_t *out) {
  FAPG_CHECK(h);
  std::vector<int32_t> v = h->g->getSubGraphPath(id);
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "path");
}

int fapg_get_all_path(fapg *h, int32_t *out) {
  FAPG_CHECK(h);
  std::vector<int32_t> v = h->g->getAllPath();
  if (v.empty() && h->g->status() != 0) return from_status(h->g);
  return copy_out(v, out, "path");
}

int fapg_get_path_from_one_point(fapg *h, int32_t v, int32_t *out) {
  FAPG_CHECK(h);
  std::vector<int32_t> r = h->g->getPathFromOnePoint(v);
  if (r.empty()) return from_status(h->g);
  return copy_out(r, out, "row");
}

int fapg_path_p2p(fapg *h, int32_t src, int32_t dst, int32_t *out, int32_t cap, int32_t *len) {
  FAPG_CHECK(h);
  if (!len || cap < 0 || (cap > 0 && !out)) return fail(FAP_E_ARG, "bad output");
  std::vector<int32_t> seq = h->g->getPathP2P(src, dst);
  if (h->g->status() != 0) return from_status(h->g);
  *len = (int32_t)seq.size();
  if ((int32_t)seq.size() > cap) return fail(FAP_E_ARG, "output too small for the path");
  if (!seq.empty()) memcpy(out, seq.data(), seq.size() * sizeof(int32_t));
  return FAP_OK;
}

int fapg_subgraph_checksum(fapg *h, int32_t id, uint64_t *sum, uint64_t *weighted) {
  FAPG_CHECK(h);
  if (!sum || !weighted) return fail(FAP_E_ARG, "null output");
  if (h->g->subGraphChecksum(id, sum, weighted) != 0) return from_status(h->g);
  return FAP_OK;
}

int fapg_check_result(fapg *h, int32_t id, int *ok) {
  FAPG_CHECK(h);
  if (!ok) return fail(FAP_E_ARG, "null output");
  fap::fapGraph &g = *h->g;
  std::vector<float> d = g.getSubGraphDistance(id);
  if (d.empty()) return from_status(h->g);
  std::vector<int32_t> src = g.getSubGraphIndex(id);
  std::vector<int32_t> gid = g.getGraphId();
  *ok = g.check_result(d.data(), nullptr, src.data(), (int)src.size(), g.num_vertexs,
                       g.adj_size.data(), g.row_offset.data(), g.col_val.data(),
                       g.weight.data(), gid.data())
            ? 1
            : 0;
  return FAP_OK;
}

}  // extern "C"
