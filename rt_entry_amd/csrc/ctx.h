// Context object behind the C ABI (shared by the API translation units).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rt_entry.h"

namespace rte {
struct Mlp;
}

// Owned device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // grows only; contents are not preserved
  hipError_t reserve(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = n;
    return hipSuccess;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

struct rte_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;

  // search space
  const float* base = nullptr;
  int64_t np = 0;
  int d = 0;
  int64_t stride = 0;
  DevBuf base_own;

  // graph
  const int32_t* graph = nullptr;
  int64_t graph_np = 0;
  int degree = 0;
  int shift = 0;
  DevBuf graph_own;

  // projection
  DevBuf rot;   // d_in x d_out compact
  DevBuf bias;  // d_out: mu R
  int pca_d = 0, pca_out = 0;

  // rerank: full-precision vectors (original space)
  const float* rr_base = nullptr;
  int64_t rr_np = 0, rr_stride = 0;
  int rr_d = 0;
  DevBuf rr_own;

  // leaves
  DevBuf nodes;
  int n_nodes = 0;
  const int32_t* pids = nullptr;
  DevBuf pids_own;
  int n_leaf = 0, cap = 0;

  // staging for host-pointer calls
  DevBuf s_in, s_out, s_aux, s_aux2, s_aux3;

  // batched search: per-query shapes, query order, work counter, and the
  // captured launch (reused while its key is unchanged)
  DevBuf b_k, b_beam, b_order, b_counter, b_kin, b_bin;
  // RTE_ORDER_LEAF: key histogram (zero between launches), scatter cursors,
  // per-range work counters
  DevBuf b_hist, b_cursor, b_seg;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;

  rte::Mlp* mlp = nullptr;
  // rte_predict_entry: kd-tree leaves (the fallback), selected in the hidden
  // layers' launch
  DevBuf kd_hits;
  ~rte_ctx();
};

// record a failure message on the context and return code
int rte_fail(rte_ctx* c, int code, const char* fmt, ...);

namespace rte {
// One device array holding every buffer of a batched launch, for calls that
// must not allocate.  Layout (int32 words): counter block (16 x int64) |
// range counters (16 x 8 x int64) | key histogram (max_keys) | key cursors
// (max_keys) | k (max_queries) | beam (max_queries) | order (max_queries).
constexpr int kScratchCounterWords = 2 * 16 + 2 * 16 * 8;
struct BatchScratch {
  int32_t* base;  // device address
  int64_t max_queries;
  int max_keys;
};
constexpr int64_t batch_scratch_words(int64_t max_queries, int max_keys) {
  return kScratchCounterWords + 2 * (int64_t)max_keys + 3 * max_queries;
}
}  // namespace rte

// rte_search_batched on device buffers with every launch buffer taken from
// sc: no allocation and no captured graph (direct launches; the buffers that
// must start at zero are cleared by a kernel first).
int rte_search_batched_in(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride,
                          const int32_t* leaf, const int32_t* k_per_q, const int32_t* beam_per_q,
                          const rte_params* p, int k_stride, int32_t* ids_out, float* dists_out,
                          const rte::BatchScratch& sc);
