// preprocess.cpp -- boundary detection and the boundary-first vertex order.
//
// findBoundry (fap/utils/util_preprocess.h:32-58): a vertex is a boundary
// vertex when it is an endpoint of an arc whose endpoints lie in different
// parts (only vertices of parts 1..K are scanned, both endpoints marked).
// sort_and_encode (:60-106): the vertices 0..V-1, each tagged (part,
// boundary flag), are sorted with std::sort by (part ascending, boundary
// first).  std::sort is not stable, so the order inside a (part, flag) class
// is whatever libstdc++'s introsort makes of the input sequence; the same
// element layout, comparator and input sequence are used here so that st2ed
// (and with it the row order of every result block) is the reference's.
// Offsets are cumulative part sizes; the boundary offset equals the part
// offset because boundary vertices come first in each part.
#include <algorithm>
#include <vector>

#include "host.h"

namespace fapi {

namespace {
struct Tagged {
  int v;
  int part;
  int bdy;
};
bool tagged_less(const Tagged &a, const Tagged &b) {
  if (a.part == b.part) return a.bdy > b.bdy;
  return a.part < b.part;
}
}  // namespace

void precondition(const HostCSR &g, int32_t K, const std::vector<int32_t> &gid, PartInfo *out) {
  const int32_t V = g.V;
  out->is_bdy.assign(V, 0);
  for (int32_t u = 0; u < V; u++) {
    const int32_t pu = gid[u];
    if (pu < 1 || pu > K) continue;
    for (int32_t e = g.row_offset[u]; e < g.row_offset[u + 1]; e++) {
      const int32_t v = g.col[e];
      if (gid[v] != pu) {
        out->is_bdy[u] = 1;
        out->is_bdy[v] = 1;
      }
    }
  }
  std::vector<Tagged> t(V);
  for (int32_t i = 0; i < V; i++) t[i] = Tagged{i, gid[i], out->is_bdy[i]};
  std::sort(t.begin(), t.end(), tagged_less);
  out->st2ed.resize(V);
  out->ed2st.resize(V);
  for (int32_t i = 0; i < V; i++) {
    out->st2ed[i] = t[i].v;
    out->ed2st[t[i].v] = i;
  }
  out->ver_num.assign(K + 1, 0);
  out->ver_off.assign(K + 1, 0);
  out->bdy_num.assign(K + 1, 0);
  out->bdy_off.assign(K + 1, 0);
  for (int32_t i = 0; i < V; i++) {
    if (gid[i] < 1 || gid[i] > K) continue;
    out->ver_num[gid[i]]++;
    if (out->is_bdy[i]) out->bdy_num[gid[i]]++;
  }
  int32_t cnt = 0;
  for (int32_t p = 1; p <= K; p++) {
    out->ver_off[p] = cnt;
    out->bdy_off[p] = cnt;
    cnt += out->ver_num[p];
  }
}

}  // namespace fapi
