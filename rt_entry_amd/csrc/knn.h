// Exact k-nearest-neighbour lists and the search graph built from them (knn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {

constexpr int kKnnK = 64;        // list length computed by the kernel (k <= 64)
constexpr int kKnnMaxD = 256;    // dims held in registers

// ids/dists: nq x k (k <= kKnnK), the k nearest rows of X (np x d, row stride
// `stride`) to each of rows [row0, row0 + nq) of X, excluding the row itself,
// ordered by (squared distance, id).
hipError_t launch_knn(const float* X, int64_t np, int d, int64_t stride, const float* norms,
                      int64_t row0, int64_t nq, int k, int32_t* ids, float* dists,
                      hipStream_t st);
// Search graph from kNN lists (np x kf, by (dist, id)): forward and reverse
// (ranks < rev_rank) candidates, duplicates keep their smallest (rank, type),
// each row keeps `degree` by (rank, type, id), padded with np.  `work` holds
// graph_from_knn_work_bytes() bytes.
hipError_t launch_graph_from_knn(const int32_t* knn, int64_t np, int kf, int degree, int rev_rank,
                                 int32_t* graph, void* work, hipStream_t st);
size_t graph_from_knn_work_bytes(int64_t np, int kf, int rev_rank);
hipError_t launch_row_norms(const float* X, int64_t np, int d, int64_t stride, float* norms,
                            hipStream_t st);

}  // namespace rte
