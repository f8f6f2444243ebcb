// Host C++ API of the query path over the C ABI (see graph.h).
#include "graph.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <numeric>
#include <sstream>

namespace rte_host {

namespace {

bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

int64_t file_size(std::ifstream& f) {
  f.seekg(0, std::ios::end);
  const int64_t n = f.tellg();
  f.seekg(0, std::ios::beg);
  return n;
}

std::ifstream open_in(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw Error("cannot open " + path);
  return f;
}

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// device buffer owned by the host API
struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

// Text rows (file_read.cu:103-119): "n d", then n*d floats, each followed
// by an optional comma; whitespace anywhere between them.
void read_txt(const std::string& path, int& n, int& d, std::vector<float>& out) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) throw Error("cannot open " + path);
  int nn = 0, dd = 0;
  if (fscanf(f, "%d %d", &nn, &dd) != 2 || nn <= 0 || dd <= 0) {
    fclose(f);
    throw Error("bad txt header in " + path);
  }
  out.resize((size_t)nn * dd);
  for (size_t i = 0; i < out.size(); ++i) {
    if (fscanf(f, " %f", &out[i]) != 1) {
      fclose(f);
      throw Error("short txt data in " + path);
    }
    int c = fgetc(f);  // the separator, if it is a comma
    if (c != ',' && c != EOF) ungetc(c, f);
  }
  fclose(f);
  n = nn;
  d = dd;
}

// By extension, in the reference's order (file_read.cu:16-80): txt, fvecs
// (:121-145), fbin (:147-161), hvecs (:163-182: int32 d then floats,
// n = (bytes/4 - 1)/d), bvecs (:184-211, at most 100M rows), hh (:82-101:
// int32 d then floats to EOF).
void read_data(const std::string& path, int& n, int& d, std::vector<float>& out) {
  if (ends_with(path, "txt")) {
    read_txt(path, n, d, out);
    return;
  }
  std::ifstream f = open_in(path);
  const int64_t size = file_size(f);
  if (ends_with(path, "fvecs")) {
    int32_t dd = 0;
    f.read((char*)&dd, 4);
    if (dd <= 0) throw Error("bad fvecs header in " + path);
    d = dd;
    n = (int)(size / (4LL * (d + 1)));
    out.resize((size_t)n * d);
    f.seekg(0);
    for (int i = 0; i < n; ++i) {
      f.seekg(4, std::ios::cur);
      f.read((char*)&out[(size_t)i * d], 4LL * d);
    }
  } else if (ends_with(path, "fbin")) {
    int32_t hdr[2];
    f.read((char*)hdr, 8);
    n = hdr[0];
    d = hdr[1];
    if (n <= 0 || d <= 0 || 8 + 4LL * n * d > size) throw Error("bad fbin header in " + path);
    out.resize((size_t)n * d);
    f.read((char*)out.data(), 4LL * n * d);
  } else if (ends_with(path, "hvecs")) {
    int32_t dd = 0;
    f.read((char*)&dd, 4);
    if (dd <= 0) throw Error("bad hvecs header in " + path);
    d = dd;
    n = (int)((size / 4 - 1) / d);
    out.resize((size_t)n * d);
    f.read((char*)out.data(), 4LL * n * d);
  } else if (ends_with(path, "bvecs")) {
    int32_t dd = 0;
    f.read((char*)&dd, 4);
    if (dd <= 0) throw Error("bad bvecs header in " + path);
    d = dd;
    n = (int)std::min<int64_t>(size / (4 + d), 100000000);
    out.resize((size_t)n * d);
    std::vector<uint8_t> row(d);
    f.seekg(0);
    for (int i = 0; i < n; ++i) {
      f.seekg(4, std::ios::cur);
      f.read((char*)row.data(), d);
      for (int j = 0; j < d; ++j) out[(size_t)i * d + j] = row[j];
    }
  } else if (ends_with(path, "hh")) {
    int32_t dd = 0;
    f.read((char*)&dd, 4);
    d = dd;
    if (d <= 0) throw Error("bad hh header in " + path);
    n = (int)((size - 4) / (4LL * d));
    out.resize((size_t)n * d);
    f.read((char*)out.data(), 4LL * n * d);
  } else {
    throw Error("unknown data file type: " + path);
  }
  if (!f) throw Error("short read in " + path);
}

// file_read.cu:213-235
void read_ivecs(const std::string& path, int& n, int& d, std::vector<int32_t>& out) {
  std::ifstream f = open_in(path);
  const int64_t size = file_size(f);
  int32_t dd = 0;
  f.read((char*)&dd, 4);
  if (dd <= 0) throw Error("bad ivecs header in " + path);
  d = dd;
  n = (int)(size / (4LL * (d + 1)));
  out.resize((size_t)n * d);
  f.seekg(0);
  for (int i = 0; i < n; ++i) {
    f.seekg(4, std::ios::cur);
    f.read((char*)&out[(size_t)i * d], 4LL * d);
  }
  if (!f) throw Error("short read in " + path);
}

// file_read.cu:237-255: raw int32 rows, degree = file size / (4 n)
void read_graph(const std::string& path, int n, int& degree, std::vector<int32_t>& out) {
  std::ifstream f = open_in(path);
  const int64_t size = file_size(f);
  degree = (int)(size / (4LL * n));
  if (degree <= 0) throw Error("graph file too short: " + path);
  out.resize((size_t)n * degree);
  f.read((char*)out.data(), 4LL * n * degree);
  if (!f) throw Error("short read in " + path);
}

// pca.cpp:70-89, 128-139: uint32 n, uint32 d, then n*d floats
void write_fbin(const std::string& path, uint32_t n, uint32_t d, const float* data) {
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) throw Error("cannot write " + path);
  f.write((const char*)&n, 4);
  f.write((const char*)&d, 4);
  f.write((const char*)data, 4LL * n * d);
}

Graph::Graph(int /*n_subspaces*/, int /*buffer_size*/, int n_candidates, int /*max_hits*/,
             double expand_ratio, double point_ratio, std::string data_name,
             std::string& data_path, std::string& query_path, std::string& gt_path,
             std::string& graph_path, int ALGO, int search_width, int topk, int max_iter,
             int device, int mode)
    : n_candidates_(n_candidates),
      topk_(topk),
      max_iter_(max_iter),
      search_width_(search_width),
      algo_(ALGO),
      mode_(mode),
      expand_ratio_((float)expand_ratio),
      point_ratio_((float)point_ratio),
      data_name_(std::move(data_name)),
      data_path_(data_path),
      query_path_(query_path),
      gt_path_(gt_path),
      graph_path_(graph_path) {
  // graph.cu:53-54: beam rounded up to a power of two; topk 0 means beam
  int nc = 1;
  while (nc < n_candidates_) nc <<= 1;
  n_candidates_ = nc;
  if (topk_ == 0) topk_ = n_candidates_;
  check(rte_create(&ctx_, device), "rte_create");
}

Graph::~Graph() { CleanUp(); }

void Graph::check(int rc, const char* what) {
  if (rc != RTE_OK) {
    std::ostringstream os;
    os << what << " failed (" << rc << "): " << (ctx_ ? rte_last_error(ctx_) : "");
    throw Error(os.str());
  }
}

void Graph::log_line(const std::string& s) {
  printf("%s", s.c_str());
  std::ofstream f(out_file_, std::ios::app);
  f << s << std::flush;
}

void Graph::Input() {  // graph.cu:73-112
  read_data(data_path_, np_, dim_, points_);
  int qd = 0;
  read_data(query_path_, nq_, qd, queries_);
  if (qd != dim_) throw Error("query and base dimensions differ");
  int gt_n = 0;
  read_ivecs(gt_path_, gt_n, gt_k_, gt_);
  // graph_helper.cu:64-121 reads the first nq rows; a shorter file cannot
  // score every query
  if (gt_n < nq_)
    throw Error("ground truth has " + std::to_string(gt_n) + " rows for " +
                std::to_string(nq_) + " queries");
  if (search_dim_ <= 0 || search_dim_ > dim_) search_dim_ = dim_;
}

void Graph::RB_Graph() {  // graph.cu:114-139
  if (!exists(graph_path_)) throw Error("graph file missing (graph building is not part of the path): " + graph_path_);
  read_graph(graph_path_, np_, degree_, graph_);
  check(rte_set_graph(ctx_, graph_.data(), np_, degree_, RTE_HOST), "rte_set_graph");
}

void Graph::Projection() {  // graph.cu:141-252, pca.cpp
  if (algo_ == 0) {  // raw space (graph.cu:306-310)
    check(rte_set_base(ctx_, points_.data(), np_, dim_, dim_, RTE_HOST), "rte_set_base");
    return;
  }
  const std::string rot = data_name_ + "/O.fbin", mean = data_name_ + "/mean.fbin",
                    pb = data_name_ + "/pca_base.fbin";
  if (exists(rot) && exists(mean)) {
    int n1, d1;
    read_data(rot, n1, d1, R_);
    if (n1 != dim_ || d1 != dim_) throw Error("rotation shape mismatch in " + rot);
    read_data(mean, n1, d1, mu_);
    if (n1 != 1 || d1 != dim_) throw Error("mean shape mismatch in " + mean);
  } else {
    R_.resize((size_t)dim_ * dim_);
    mu_.resize(dim_);
    check(rte_pca_fit(ctx_, points_.data(), np_, dim_, dim_, R_.data(), mu_.data(), nullptr,
                      RTE_HOST),
          "rte_pca_fit");
    mkdir(data_name_.c_str(), 0755);
    write_fbin(rot, dim_, dim_, R_.data());
    write_fbin(mean, 1, dim_, mu_.data());
  }
  if (exists(pb)) {
    int n1, d1;
    read_data(pb, n1, d1, pca_points_);
    if (n1 != np_ || d1 != dim_) throw Error("pca_base shape mismatch in " + pb);
  } else {
    pca_points_.resize((size_t)np_ * dim_);
    check(rte_pca_rotate(ctx_, points_.data(), np_, dim_, dim_, R_.data(), mu_.data(), dim_,
                         pca_points_.data(), RTE_HOST),
          "rte_pca_rotate");
    write_fbin(pb, np_, dim_, pca_points_.data());
  }
  // search space: the first search_dim_ PCA columns, row stride dim_
  check(rte_set_base(ctx_, pca_points_.data(), np_, search_dim_, dim_, RTE_HOST), "rte_set_base");
  check(rte_set_pca(ctx_, R_.data(), mu_.data(), dim_, dim_, search_dim_, RTE_HOST), "rte_set_pca");
}

void Graph::Init_entry() {  // entry.cu:8-98
  if (algo_ != 1) return;
  const int cap = (int)((float)np_ * point_ratio_);  // entry.cu:22
  if (cap <= 0) throw Error("point_ratio gives an empty leaf capacity");
  const std::string cache = data_name_ + "/" + std::to_string(cap) + "_0.aabbs";
  std::vector<float> boxes;
  std::vector<int32_t> pids;
  int n_leaf = 0;
  if (exists(cache)) {  // entry.cu:183-190
    std::ifstream f = open_in(cache);
    int32_t hdr[2];
    f.read((char*)hdr, 8);
    n_leaf = hdr[0];
    if (hdr[1] != cap || n_leaf <= 0) throw Error("aabb cache shape mismatch in " + cache);
    boxes.resize((size_t)n_leaf * 6);
    pids.resize((size_t)n_leaf * cap);
    f.read((char*)boxes.data(), 24LL * n_leaf);
    f.read((char*)pids.data(), 4LL * n_leaf * cap);
    if (!f) throw Error("short read in " + cache);
  } else {
    const int max_leaves = 4 * (np_ / cap) + 16;
    boxes.resize((size_t)max_leaves * 6);
    pids.resize((size_t)max_leaves * cap);
    if (cap <= 2048) {  // GPU build over the first three PCA columns
      check(rte_build_kdtree_gpu(ctx_, pca_points_.data(), np_, dim_, cap, max_leaves,
                                 boxes.data(), pids.data(), &n_leaf, RTE_HOST),
            "rte_build_kdtree_gpu");
    } else {
      std::vector<float> pts3((size_t)np_ * 3);
      for (int64_t i = 0; i < np_; ++i)
        for (int a = 0; a < 3; ++a) pts3[i * 3 + a] = pca_points_[i * dim_ + a];
      check(rte_build_kdtree(pts3.data(), np_, cap, max_leaves, boxes.data(), pids.data(),
                             &n_leaf),
            "rte_build_kdtree");
    }
    boxes.resize((size_t)n_leaf * 6);
    pids.resize((size_t)n_leaf * cap);
    mkdir(data_name_.c_str(), 0755);
    std::ofstream f(cache, std::ios::binary);  // entry.cu:171-180
    const int32_t hdr[2] = {n_leaf, cap};
    f.write((const char*)hdr, 8);
    f.write((const char*)boxes.data(), 24LL * n_leaf);
    f.write((const char*)pids.data(), 4LL * n_leaf * cap);
  }
  log_line("n_aabbs: " + std::to_string(n_leaf) + "\n");  // entry.cu:55-58
  rte_expand_aabbs(boxes.data(), n_leaf, expand_ratio_);
  check(rte_set_leaves(ctx_, boxes.data(), pids.data(), n_leaf, cap, RTE_HOST), "rte_set_leaves");
}

void Graph::Search() {  // graph.cu:254-293
  const int qdim = algo_ == 0 ? dim_ : search_dim_;
  int beam = 1;
  while (beam < n_candidates_) beam <<= 1;  // graph.cu:53
  const int kk = reorder_ ? beam : topk_;  // REORDER: the whole beam (graph_search.cuh:1955)
  if (reorder_)
    check(rte_set_rerank_base(ctx_, points_.data(), np_, dim_, dim_, RTE_HOST),
          "rte_set_rerank_base");
  Dev dq, dpq, dleaf, dids, ddist, dcand;
  hip_ok(hipMalloc(&dq.p, (size_t)nq_ * dim_ * 4), "hipMalloc queries");
  hip_ok(hipMemcpy(dq.p, queries_.data(), queries_.size() * 4, hipMemcpyHostToDevice),
         "copy queries");
  hip_ok(hipMalloc(&dpq.p, (size_t)nq_ * std::max(qdim, 3) * 4), "hipMalloc pq");
  hip_ok(hipMalloc(&dleaf.p, (size_t)nq_ * 4), "hipMalloc leaf");
  hip_ok(hipMalloc(&dids.p, (size_t)nq_ * topk_ * 4), "hipMalloc ids");
  hip_ok(hipMalloc(&ddist.p, (size_t)nq_ * topk_ * 4), "hipMalloc dists");
  if (reorder_) hip_ok(hipMalloc(&dcand.p, (size_t)nq_ * kk * 4), "hipMalloc candidates");
  rte_params p{};
  p.k = kk;
  p.beam = n_candidates_;
  p.width = search_width_;
  p.max_iter = max_iter_;
  p.mode = mode_;
  p.algo = algo_;
  check(rte_synchronize(ctx_), "sync");
  const auto t0 = std::chrono::high_resolution_clock::now();
  const float* sq = (const float*)dq.p;
  if (algo_ != 0) {
    check(rte_project(ctx_, (const float*)dq.p, nq_, (float*)dpq.p, RTE_DEVICE), "rte_project");
    sq = (const float*)dpq.p;
  }
  if (algo_ == 1)
    check(rte_select_entry(ctx_, sq, nq_, qdim, (int32_t*)dleaf.p, RTE_DEVICE),
          "rte_select_entry");
  check(rte_search(ctx_, sq, nq_, qdim, (const int32_t*)dleaf.p, &p,
                   (int32_t*)(reorder_ ? dcand.p : dids.p), reorder_ ? nullptr : (float*)ddist.p,
                   nullptr, RTE_DEVICE),
        "rte_search");
  if (reorder_) {  // graph.cu:326-336, on the device; inside the search timer
    check(rte_synchronize(ctx_), "sync");
    const auto t2 = std::chrono::high_resolution_clock::now();
    check(rte_rerank(ctx_, (const float*)dq.p, nq_, dim_, (const int32_t*)dcand.p, kk, topk_,
                     (int32_t*)dids.p, (float*)ddist.p, RTE_DEVICE),
          "rte_rerank");
    check(rte_synchronize(ctx_), "sync");
    reorder_ms_ = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t2)
                      .count() * 1000.0;
    std::ostringstream ro;  // the nested timer stops first (Timing.h:41-46)
    ro << "time reorder: " << reorder_ms_ << " ms\n";
    log_line(ro.str());
  }
  check(rte_synchronize(ctx_), "sync");
  const auto t1 = std::chrono::high_resolution_clock::now();
  // graph.cu:255-286: projection + entry + graph search (+ REORDER)
  search_ms_ = std::chrono::duration<double>(t1 - t0).count() * 1000.0;
  std::ostringstream os;
  os << "time search: " << search_ms_ << " ms\n";  // Timing.h:41-46
  log_line(os.str());
  results_.resize((size_t)nq_ * topk_);
  dists_.resize((size_t)nq_ * topk_);
  hits_.resize(nq_);
  hip_ok(hipMemcpy(results_.data(), dids.p, results_.size() * 4, hipMemcpyDeviceToHost), "ids");
  hip_ok(hipMemcpy(dists_.data(), ddist.p, dists_.size() * 4, hipMemcpyDeviceToHost), "dists");
  if (algo_ == 1)
    hip_ok(hipMemcpy(hits_.data(), dleaf.p, hits_.size() * 4, hipMemcpyDeviceToHost), "leaves");
  // graph_helper.cu:64-121: true top-10 found among the topk results
  int64_t found = 0;
  const int g10 = std::min(10, gt_k_);
  for (int q = 0; q < nq_; ++q)
    for (int i = 0; i < g10; ++i) {
      const int32_t g = gt_[(size_t)q * gt_k_ + i];
      for (int j = 0; j < topk_; ++j)
        if (results_[(size_t)q * topk_ + j] == g) {
          ++found;
          break;
        }
    }
  recall10_ = (double)found / nq_ / 10.0;
  std::ostringstream rs;
  rs << "recall@10 = " << recall10_ << " ms\n";  // graph_helper.cu:118, suffix as written
  log_line(rs.str());
}

void Graph::CleanUp() {
  if (ctx_) {
    rte_destroy(ctx_);
    ctx_ = nullptr;
  }
}

}  // namespace rte_host
