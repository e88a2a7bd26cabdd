// Host check of the boundary index (fastapsp_amd/csrc/bindex.hip): its host
// code -- index build, tile and border lists, the block iterations'
// workspace layout and bookkeeping -- compiled as it is (the test slices it
// out of the source) against a host-only stand-in of the HIP runtime
// (mock/hip/hip_runtime.h) and CPU emulations of its kernels
// (emu_kernels.h), under AddressSanitizer / UBSan.  Reads a partitioned
// graph written by tests/test_host.py, solves the boundary rows of every
// group of parts and compares sampled rows with Dijkstra.
// Usage: bidx_host <graph.bin> <parts per call> [aliased row buffers] [max calls] [band]
// Prints "checked N rows, 0 bad" and exits 0 when everything matches.
#include <stdio.h>
#include <math.h>
#include <numeric>
#include <queue>
#include <string>
#include <vector>
#include "common.h"
namespace fapi {
void set_error(const std::string &m) { fprintf(stderr, "error: %s\n", m.c_str()); }
void clear_error() {}
int fail(int code, const std::string &m) { set_error(m); return code; }
int hip_fail(hipError_t, const char *w, const char *f, int l) { fprintf(stderr, "hip %s %s:%d\n", w, f, l); return FAP_E_HIP; }
int ensure_ws(fap_ctx *, Workspace &ws, size_t bytes) {
  if (ws.bytes >= bytes) return FAP_OK;
  free(ws.ptr);
  ws.ptr = malloc(bytes);
  ws.bytes = bytes;
  return FAP_OK;
}
int launch_fill_i32(int32_t *d, int64_t n, int32_t v, hipStream_t) { for (int64_t i = 0; i < n; i++) d[i] = v; return FAP_OK; }
int launch_gather(const fap_graph_dev *g, int32_t start, int32_t n, int32_t n_bdy, const int32_t *, int64_t,
                  int32_t *sub, int64_t ld, hipStream_t) {
  if (n_bdy != 0) abort();
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) sub[(int64_t)i * ld + j] = i == j ? 0 : kInf;
    const int u = g->st2ed[start + i];
    for (int e = g->row_offset[u]; e < g->row_offset[u + 1]; e++) {
      const int j = g->ed2st[g->edges[e].x] - start;
      if (j >= 0 && j < n && j != i) sub[(int64_t)i * ld + j] = g->edges[e].y;
    }
  }
  return FAP_OK;
}
}  // namespace fapi
extern "C" int fap_fw_batched(fap_ctx *, int32_t count, int32_t *const *m, const int32_t *ns, const int64_t *lds, void *);
#include "bindex_host.inc"

static std::vector<int32_t> dijkstra(const std::vector<int32_t> &ro, const std::vector<int2> &ed, int s) {
  std::vector<int32_t> d(ro.size() - 1, fapi::kInf);
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<>> pq;
  d[s] = 0; pq.push({0, s});
  while (!pq.empty()) {
    auto [du, u] = pq.top(); pq.pop();
    if (du != d[u]) continue;
    for (int e = ro[u]; e < ro[u + 1]; e++) {
      const int v = ed[e].x, nd = du + ed[e].y;
      if (nd < d[v]) { d[v] = nd; pq.push({nd, v}); }
    }
  }
  return d;
}

int main(int argc, char **argv) {
  setvbuf(stdout, nullptr, _IONBF, 0);
  FILE *f = fopen(argv[1], "rb");
  const int max_groups = argc > 4 ? atoi(argv[4]) : 1 << 30;
  int ngroups = 0;
  const int group = atoi(argv[2]), alias = argc > 3 ? atoi(argv[3]) : 0;
  int64_t hdr[3];
  fread(hdr, 8, 3, f);
  const int V = hdr[0], E = hdr[1], K = hdr[2];
  auto rd = [&](size_t n) { std::vector<int32_t> v(n); fread(v.data(), 4, n, f); return v; };
  auto ro = rd(V + 1), col = rd(E), wt = rd(E), gid = rd(V), st2ed = rd(V), ed2st = rd(V);
  auto vn = rd(K + 1), vo = rd(K + 1), bn = rd(K + 1);
  fclose(f);
  std::vector<int2> ed(E);
  for (int e = 0; e < E; e++) ed[e] = int2{col[e], wt[e]};
  fap_graph_dev g;
  g.V = V; g.E = E;
  g.row_offset = ro.data(); g.edges = ed.data(); g.graph_id = gid.data(); g.st2ed = st2ed.data(); g.ed2st = ed2st.data();
  g.h_st2ed = st2ed; g.h_graph_id = gid;
  fap_ctx *ctx = new fap_ctx();
  ctx->num_cus = 256;
  if (argc > 5) ctx->opt_bidx_delta = atoi(argv[5]);
  int rc = fapi::bidx_build(ctx, &g);
  printf("build rc=%d bidx=%p delta=%d nb=%d nbp=%d bytes=%zu\n", rc, (void *)g.bidx, g.bidx ? g.bidx->delta : 0, g.bidx ? g.bidx->nb : 0,
         g.bidx ? g.bidx->nbp : 0, g.bidx ? g.bidx->bytes : 0);
  if (!g.bidx) return 1;
  long checked = 0, bad = 0;
  for (int p0 = 1; p0 <= K && ngroups < max_groups; p0 += group, ngroups++) {
    const int cnt = std::min(group, K + 1 - p0);
    std::vector<int32_t> qidx(cnt), rowoff(cnt);
    int nsrc = 0;
    for (int i = 0; i < cnt; i++) {
      const int p = p0 + i;
      qidx[i] = fapi::bidx_part(&g, vo[p], vn[p], bn[p]);
      if (qidx[i] < 0) { printf("part %d not held\n", p); return 1; }
      rowoff[i] = nsrc;
      nsrc += bn[p];
    }
    const int nbuf = alias ? std::min(alias, nsrc) : nsrc;
    std::vector<int32_t> out((size_t)nbuf * V, -7);
    std::vector<int32_t *> rp(nsrc);
    for (int r = 0; r < nsrc; r++) rp[r] = out.data() + (size_t)(r % nbuf) * V;
    fapk::g_tiles = fapk::g_launches = fapk::g_max_tiles = fapk::g_kwork = 0;
    rc = fapi::bidx_rows(ctx, &g, cnt, qidx.data(), rowoff.data(), nsrc, rp.data(), nullptr);
    printf("parts %d..%d nsrc %d rc %d: %ld launches, %ld tiles, max %ld per launch, mean K %.1f\n", p0, p0 + cnt - 1,
           nsrc, rc, fapk::g_launches, fapk::g_tiles, fapk::g_max_tiles, (double)fapk::g_kwork / std::max(1L, fapk::g_tiles));
    fflush(stdout);
    if (rc) return 1;
    if (alias) continue;
    for (int i = 0, r = 0; i < cnt; i++)
      for (int j = 0; j < bn[p0 + i]; j++, r++) {
        if (j % 17 && j != bn[p0 + i] - 1) continue;
        const auto want = dijkstra(ro, ed, st2ed[vo[p0 + i] + j]);
        checked++;
        if (memcmp(want.data(), rp[r], 4 * (size_t)V)) { bad++; if (bad < 5) printf("mismatch part %d row %d\n", p0 + i, j); }
      }
  }
  printf("checked %ld rows, %ld bad\n", checked, bad);
  return bad != 0;
}
