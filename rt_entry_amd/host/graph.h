// Host C++ API of the query path, mirroring the reference's Graph class
// (graph.h:10-82, graph.cu:32-342) over the C ABI of include/rt_entry.h.
//
// Same constructor arguments and call order as the reference's main
// (main.cpp:82-94): Input(); RB_Graph(); Projection(); Init_entry();
// Search(); CleanUp().  Differences, by design:
//  - errors throw rte_host::Error instead of calling exit();
//  - Search() keeps ids and distances (results(), distances());
//  - ALGO 1 entries come from the GPU kd-tree traversal (rte_select_entry)
//    instead of an OptiX ray cast;
//  - the artefact caches keep the reference's file names and formats
//    (<data_name>/O.fbin, mean.fbin, pca_base.fbin,
//    <aabb_size>_0.aabbs) so a reference cache directory can be reused.
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "rt_entry.h"

namespace rte_host {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// file_read.cu:16-255: fvecs / fbin / bvecs / hh by extension, ivecs, .nsw
void read_data(const std::string& path, int& n, int& d, std::vector<float>& out);
void read_ivecs(const std::string& path, int& n, int& d, std::vector<int32_t>& out);
void read_graph(const std::string& path, int n, int& degree, std::vector<int32_t>& out);
void write_fbin(const std::string& path, uint32_t n, uint32_t d, const float* data);

class Graph {
 public:
  Graph(int n_subspaces, int buffer_size, int n_candidates, int max_hits, double expand_ratio,
        double point_ratio, std::string data_name, std::string& data_path,
        std::string& query_path, std::string& gt_path, std::string& graph_path, int ALGO,
        int search_width, int topk, int max_iter, int device = 0,
        int mode = RTE_MODE_REF);
  ~Graph();
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  void Input();
  void RB_Graph();
  void Projection();
  void Init_entry();
  void Search();
  void CleanUp();

  // dimension of the search space (the reference's DIM); default: full d
  void set_search_dim(int dim) { search_dim_ = dim; }
  // REORDER (graph.cu:326-336): search returns the whole beam, reranked on
  // the original vectors to topk
  void set_reorder(bool on) { reorder_ = on; }
  void set_out_file(const std::string& path) { out_file_ = path; }

  const std::vector<int32_t>& results() const { return results_; }
  const std::vector<float>& distances() const { return dists_; }
  const std::vector<int32_t>& entries() const { return hits_; }
  double recall10() const { return recall10_; }
  double search_ms() const { return search_ms_; }
  int nq() const { return nq_; }
  int topk() const { return topk_; }

 private:
  void check(int rc, const char* what);
  void log_line(const std::string& s);

  rte_ctx* ctx_ = nullptr;
  int n_candidates_, topk_, max_iter_, search_width_, algo_, mode_;
  float expand_ratio_, point_ratio_;
  std::string data_name_, data_path_, query_path_, gt_path_, graph_path_;
  std::string out_file_ = "out.txt";
  int np_ = 0, nq_ = 0, dim_ = 0, gt_k_ = 0, degree_ = 0, search_dim_ = 0;
  std::vector<float> points_, queries_, pca_points_, R_, mu_, pca_queries_;
  std::vector<int32_t> gt_, graph_, hits_, results_;
  std::vector<float> dists_;
  double recall10_ = 0.0, search_ms_ = 0.0, reorder_ms_ = 0.0;
  bool reorder_ = false;
};

}  // namespace rte_host
