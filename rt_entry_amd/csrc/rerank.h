// Full-precision rerank of search candidates (rerank.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {

constexpr int kRerankMaxCand = 512;

hipError_t launch_rerank(const float* base, int64_t np, int64_t stride, int d, const float* Q,
                         int64_t nq, int64_t q_stride, const int32_t* cand, int n_cand, int k,
                         int32_t* ids_out, float* dists_out, hipStream_t st);

}  // namespace rte
