// fap/fap.h -- drop-in fap::fapGraph on MI355X.
//
// Same public API as the reference's fap/fap.h:22-125 (constructor from a
// .mtx name, CSR constructor, preCondition, isSplit, solve, solveSubGraph and
// the getters, public CSR members).  Behind it, every solve runs on one GPU
// through the C ABI in fap_dev.h: the CSR and the partition stay resident in
// HBM, and result blocks stay in HBM until a getter asks for them.
//
// Differences from the reference, each deliberate:
//   * int32 arithmetic (INF = FAP_INF); the float getters convert exactly
//     (INF -> 1e8f, the reference's MAXVALUE), int32 twins are added.
//   * solve / solveSubGraph return 0 on success and -1 on error, as the
//     reference's header comments say (its bodies return nothing); the reason
//     is in lastError().  Errors never exit the process.
//   * the CSR constructor (a stub in the reference, fap/fap.cpp:61-65) works.
//   * getDistanceFromOnePoint / getPathFromOnePoint index the current part's
//     rows (the reference indexes by vertex id, fap/fap.cpp:247-248).
//   * solve() leaves getCurrentSubGraphId() at K (the reference never sets
//     it) and getSubGraphDistance(id) then returns part id's rows of the full
//     matrix.
//   * path matrices (is_path_needed) hold, per row, the vertex before each
//     column on a shortest path (the reference's semantics); which of several
//     equally short paths is not fixed (the reference's backends differ too).
//   * getPathP2P(src, dst) returns the vertex sequence src .. dst (the
//     no-argument reference stub, fap/fap.cpp:267-269, reports an error).
#ifndef FAP_FAP_H_
#define FAP_FAP_H_

#include <stdint.h>

#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace fap {

class fapGraph {
 protected:
  std::string input_graph;
  std::stringstream input_file;

  // parts above this many vertices are solved part by part (fap/fap.h:32)
  const int32_t max_vertexs_num_limit = 20000;
  bool is_split = false;
  bool is_path_needed = false;
  std::string output_file;
  std::vector<float> dist;
  std::vector<int32_t> path;
  int32_t current_subgraph_id = 0;
  std::vector<float> subgraph_dist;
  std::vector<int32_t> subgraph_path;

  int32_t K = 0;
  std::vector<int> graph_id;
  std::unordered_map<int, std::vector<int>> BlockVer;
  std::unordered_map<int, std::vector<int>> BlockBoundary;
  std::unordered_map<int, int> isBoundry;
  std::vector<int> st2ed;
  std::vector<int> ed2st;
  std::vector<int> C_BlockVer_num;
  std::vector<int> C_BlockVer_offset;
  std::vector<int> C_BlockBdy_num;
  std::vector<int> C_BlockBdy_offset;
  int32_t max_subgraph_vertexs = 0;

 private:
  struct Device;  // GPU-side state (context, resident graph, result blocks)
  std::unique_ptr<Device> dev_;
  std::vector<int32_t> weight_i32_;
  bool loaded_ = false;
  bool preconditioned_ = false;
  bool full_solved_ = false;
  bool used_metis_ = false;
  bool batched_ = true;
  int32_t status_ = 0;
  std::string error_;

  int32_t fail(int32_t code, const std::string &msg);
  void reset_status();
  int32_t ensure_device();
  int32_t run(int32_t subgraph_id, int32_t *d_rowblock);
  int32_t run_batch(const std::vector<int32_t> &ids, const std::vector<int32_t *> &blocks);
  int32_t run_pred(int32_t subgraph_id, const int32_t *d_rowblock, int32_t *d_predblock);
  const int32_t *subGraphPathDevice(int32_t sub_garph_id);
  void init_from_csr(int32_t V, std::vector<int32_t> &&ro, std::vector<int32_t> &&col,
                     std::vector<int32_t> &&w, bool directed_graph);
  bool part_ok(int32_t id) const;

 public:
  int32_t num_vertexs = 0;
  int32_t num_edges = 0;
  std::vector<int32_t> adj_size;
  std::vector<int32_t> row_offset;
  std::vector<int32_t> col_val;
  std::vector<float> weight;
  bool directed = false;
  bool weighted = false;

  // graph/<weight|unweight>-<directed|undirected>/<input_graph>.mtx relative
  // to the working directory, or a path ending in ".mtx"; then K-way
  // partition (fap/fap.cpp:27-59).
  fapGraph(std::string input_graph, bool directed, bool weight, int32_t K);
  // From CSR arrays (weights integral, >= 0); the graph is taken as directed
  // unless every arc has its reverse.
  fapGraph(int32_t num_vertexs, int64_t num_edges, std::vector<int32_t> row_offset,
           std::vector<int32_t> col_val, std::vector<float> weight, int32_t K);
  ~fapGraph();
  fapGraph(const fapGraph &) = delete;
  fapGraph &operator=(const fapGraph &) = delete;

  // Replace the partition computed by the constructor (1-based part ids);
  // the equivalent of the reference's DEBUG id-file path (readIdFile.h:42-68).
  int32_t setPartition(const std::vector<int32_t> &graph_id);

  void preCondition();
  bool isSplit();
  std::vector<int> getGraphId();

  int32_t solve(bool is_path_needed = false);
  int32_t solveSubGraph(int32_t sub_garph_id, bool is_path_needed = false);

  std::vector<int32_t> getMapping();
  std::vector<float> getAllDistance();
  std::vector<int32_t> getAllPath();
  int32_t getCurrentSubGraphId();
  std::vector<int32_t> getSubGraphIndex(int32_t sub_garph_id);
  std::vector<float> getSubGraphDistance(int32_t sub_garph_id);
  std::vector<int32_t> getSubGraphPath(int32_t sub_garph_id);
  std::vector<float> getDistanceFromOnePoint(int32_t vectex_id);
  std::vector<int32_t> getPathFromOnePoint(int32_t vectex_id);
  float getDistanceP2P();
  std::vector<int32_t> getPathP2P();

  bool check_result(float *dist, int *path, int *source, int source_num, int vertexs,
                    int *adj_size, int *row_offset, int *col_val, float *weight, int *graph_id);

  // ---- additions ----
  // Batched (default) or plain part solves for solve(), solveSubGraphs() and
  // solveSubGraph(): batched takes the boundary rows of every part of a
  // group from the graph's boundary index (built by the first such call)
  // and closes the group's part matrices in one batched closure; plain runs
  // each part as the reference does (SSSP from its boundary vertices, FW,
  // min-plus), one after another.  Both give the same distances.
  void setBatched(bool batched) { batched_ = batched; }
  bool isBatched() const { return batched_; }
  // Several parts at once (1-based ids); each block stays on the device and
  // the getters accept any id of the last call.
  int32_t solveSubGraphs(const std::vector<int32_t> &ids, bool is_path_needed = false);
  int32_t status() const { return status_; }
  const std::string &lastError() const { return error_; }
  bool usedMetis() const { return used_metis_; }
  int32_t numParts() const { return K; }
  const std::vector<int32_t> &weightInt32() const { return weight_i32_; }
  std::vector<int32_t> getPartVertexNum() const { return C_BlockVer_num; }
  std::vector<int32_t> getPartVertexOffset() const { return C_BlockVer_offset; }
  std::vector<int32_t> getPartBoundaryNum() const { return C_BlockBdy_num; }
  std::vector<int32_t> getBoundaryFlags() const;
  // int32 results (INF = FAP_INF)
  std::vector<int32_t> getAllDistanceInt32();
  std::vector<int32_t> getSubGraphDistanceInt32(int32_t sub_garph_id);
  // rows [row0, row0 + nrows) of a solved part's block (row r = vertex
  // st2ed[offset + r]) into out (nrows x V); 0 or -1
  int32_t getSubGraphRowsInt32(int32_t sub_garph_id, int32_t row0, int32_t nrows, int32_t *out);
  // device-resident view of a solved block: row r of part id at ptr + r*ld
  const int32_t *subGraphDistanceDevice(int32_t sub_garph_id, int64_t *ld);
  // distance between two original vertex ids, from the resident result
  // (solve(), or solveSubGraph(part of src)); -1 if not available
  int32_t getDistanceP2P(int32_t src, int32_t dst, int32_t *out);
  // vertex sequence src .. dst of a shortest path, from the resident path
  // rows of src's part (solved with is_path_needed); empty if dst is not
  // reachable (status() == 0) or on error (status() != 0)
  std::vector<int32_t> getPathP2P(int32_t src, int32_t dst);
  // wrapping sums over a solved part block on the device (position keyed)
  int32_t subGraphChecksum(int32_t sub_garph_id, uint64_t *sum, uint64_t *weighted);
};

}  // namespace fap

#endif  // FAP_FAP_H_
