// Entry-leaf selection over leaf AABBs (entry.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rte {

struct BoxNode {
  float lo[3];
  float hi[3];
  int skip;  // preorder index of the next node outside this subtree
  int leaf;  // leaf id, or -1 for an inner node
};
static_assert(sizeof(BoxNode) == 32, "BoxNode is 32 bytes");

void build_hierarchy(const float* aabbs, int n_leaf, std::vector<BoxNode>& out);
hipError_t launch_select_entry(const BoxNode* nodes, int n_nodes, const float* pq, int64_t nq,
                               int stride, int32_t* hits, hipStream_t st);

}  // namespace rte
