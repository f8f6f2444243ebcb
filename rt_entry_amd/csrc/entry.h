// Entry-leaf selection over leaf AABBs (entry.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rte {

// Node of the box hierarchy, stored in breadth-first order (the top levels
// form a prefix of the array).  A walk visits nodes in preorder through the
// links: into the first child when the query is inside an inner node's box,
// else to `skip`, the next node in preorder outside this subtree (n_nodes at
// the end).
struct BoxNode {
  float lo[3];
  float hi[3];
  int skip;  // index of the next node in preorder outside this subtree
  int link;  // leaf id (>= 0), or -(first child index) - 1 for an inner node
};
static_assert(sizeof(BoxNode) == 32, "BoxNode is 32 bytes");

void build_hierarchy(const float* aabbs, int n_leaf, std::vector<BoxNode>& out);
// nodes kept in LDS by the selection kernel (the first 11 levels)
constexpr int kTopNodes = 2047;

hipError_t launch_select_entry(const BoxNode* nodes, int n_nodes, const float* pq, int64_t nq,
                               int stride, int32_t* hits, hipStream_t st);

}  // namespace rte
