// Entry-space links of a search graph (links.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {

constexpr int kMaxLinks = 32;  // links per row
constexpr int kLinkTries = 8;  // draws per link before it is left as padding

// graph: np x degree (device, in place); pids: n_leaf x cap (device);
// leaf_of (np) and size (n_leaf) are device work arrays
hipError_t launch_entry_links(int32_t* graph, int64_t np, int degree, int r, int window,
                              uint64_t seed, const int32_t* pids, int n_leaf, int cap,
                              int32_t* leaf_of, int32_t* size, hipStream_t st);

}  // namespace rte
