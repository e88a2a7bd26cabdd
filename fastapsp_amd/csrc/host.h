// host.h -- host-side graph I/O, partitioning and preprocessing (internal).
#ifndef FAP_HOST_H_
#define FAP_HOST_H_

#include <stdint.h>

#include <string>
#include <vector>

namespace fapi {

struct HostCSR {
  int32_t V = 0;
  int64_t E = 0;
  std::vector<int32_t> row_offset;  // V+1
  std::vector<int32_t> col;         // E
  std::vector<int32_t> w;           // E, int32 >= 0
};

// Reads the reference's .mtx flavour (fap/utils/readMatFile.h:36-173): leading
// '%' lines skipped, header "n t m", then 1-based "a b" (unweighted, w = 1) or
// "a b c" (weighted, w = |c|) until the first token that does not parse; an
// undirected edge yields a->b then b->a.  Each adjacency list is the reverse
// of insertion order (head insertion, :30-34, walked at :150-162).  Weights
// must be integral and < INF.  Returns 0 or a FAP_E_* code with *err set.
int read_mtx(const std::string &path, bool directed, bool weighted, HostCSR *out,
             std::string *err);

// Reference path convention: graph/{unweight,weight}-{undirected,directed}/<name>.mtx
// relative to the working directory (readMatFile.h:43-51).  A name that
// already ends in ".mtx" is used as a path as given.
std::string reference_graph_path(const std::string &name, bool directed, bool weighted);

// K-way partition, part ids 1..K in graph_id (readIdFile.h:127-131).  Uses
// libmetis (METIS_PartGraphKway, NULL options, as Graph_decomposition
// :71-132) when it can be loaded, else the built-in recursive bisection.
// *used_metis reports which one ran.
int partition_graph(const HostCSR &g, bool directed, int32_t K, std::vector<int32_t> *graph_id,
                    bool *used_metis, std::string *err);

// Built-in partitioner (no METIS): recursive bisection; each region is
// ordered by (hop distance from a - hop distance from b) for a pseudo-
// diameter pair (a, b) found by two BFS sweeps, components one after the
// other, and cut at the size ratio of the part counts.  Deterministic.
void bisection_partition(const HostCSR &g, int32_t K, std::vector<int32_t> *graph_id);

struct PartInfo {
  std::vector<int32_t> is_bdy;    // V
  std::vector<int32_t> st2ed;     // V
  std::vector<int32_t> ed2st;     // V
  std::vector<int32_t> ver_num;   // K+1 (index 0 unused)
  std::vector<int32_t> ver_off;   // K+1
  std::vector<int32_t> bdy_num;   // K+1
  std::vector<int32_t> bdy_off;   // K+1
};

// findBoundry + sort_and_encode (fap/utils/util_preprocess.h:32-106).
void precondition(const HostCSR &g, int32_t K, const std::vector<int32_t> &graph_id,
                  PartInfo *out);

}  // namespace fapi

#endif  // FAP_HOST_H_
