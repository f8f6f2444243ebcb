// Host kd-tree build over 3-D points (kdtree.cpp).
#pragma once
#include <stdint.h>

namespace rte {

// Returns 0, -2 when more than max_leaves leaves are needed, -3 when a split
// puts every point on one side.  aabbs: max_leaves x 6; pids: max_leaves x
// max_node.
int build_kdtree(const float* pts3, int64_t np, int max_node, int max_leaves, float* aabbs,
                 int32_t* pids, int* n_leaf_out);
void expand_aabbs(float* aabbs, int n_leaf, float ratio);

}  // namespace rte
