// Device side of the entry-leaf selection (entry.hip): the stackless walk,
// shared by select_entry_kernel and the predictor's fused kernel
// (mlp.hip), which runs the selection in workgroups of its own launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "entry.h"

namespace rte {

// One step of the walk over node (a, b) = {lo[3], hi[3], skip, link}: a leaf
// whose box holds the point ends the walk (hit = leaf id); an inner node
// whose box holds it continues with its first child; otherwise the walk
// continues at the skip link (later in preorder).
__device__ __forceinline__ bool visit_node(uint4 a, uint4 b, float x, float y, float z, int& i,
                                           int& hit) {
  const bool in = (x >= __uint_as_float(a.x)) & (x <= __uint_as_float(a.w)) &
                  (y >= __uint_as_float(a.y)) & (y <= __uint_as_float(b.x)) &
                  (z >= __uint_as_float(a.z)) & (z <= __uint_as_float(b.y));
  const int skip = (int)b.z, link = (int)b.w;
  if (in && link >= 0) {
    hit = link;
    return true;
  }
  i = in ? -link - 1 : skip;
  return false;
}

// Entry leaves of queries q0 + threadIdx.x (one workgroup of 256 threads):
// the top kTopNodes nodes are copied into `top` (LDS, 16-byte aligned,
// kTopNodes * 32 bytes), then each lane walks for its query.
__device__ __forceinline__ void select_entry_block(const BoxNode* __restrict__ nodes, int n_nodes,
                                                   BoxNode* top, const float* __restrict__ pq,
                                                   int64_t nq, int stride,
                                                   int32_t* __restrict__ hits, int64_t q0) {
  const int n_top = min(n_nodes, kTopNodes);
  {  // all loads of the copy first (one memory round trip), then the stores
    constexpr int kPer = (kTopNodes * 2 + 255) / 256;  // 16-byte chunks per thread
    const uint4* src = reinterpret_cast<const uint4*>(nodes);
    uint4* dst = reinterpret_cast<uint4*>(top);
    uint4 v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int i = (int)threadIdx.x + 256 * j;
      v[j] = i < 2 * n_top ? src[i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int i = (int)threadIdx.x + 256 * j;
      if (i < 2 * n_top) dst[i] = v[j];
    }
  }
  __syncthreads();
  const int64_t q = q0 + threadIdx.x;
  if (q >= nq) return;
  const float x = pq[q * stride + 0];
  const float y = pq[q * stride + 1];
  const float z = pq[q * stride + 2];
  int hit = 0;
  int i = 0;
  bool found = false;
  // Two loops, one per address space: a node index that may point into LDS
  // or global memory would make every read a generic (flat) load.  Each
  // visit reads the whole 32-byte node as two 16-byte loads and tests the box
  // without short-circuits, so a visit is one memory access.
  const uint4* top4 = reinterpret_cast<const uint4*>(top);
  const uint4* all4 = reinterpret_cast<const uint4*>(nodes);
  while (!found && i < n_nodes) {
    while (!found && i < n_top) {
      found = visit_node(top4[2 * i], top4[2 * i + 1], x, y, z, i, hit);
    }
    while (!found && i >= n_top && i < n_nodes) {
      found = visit_node(all4[2 * i], all4[2 * i + 1], x, y, z, i, hit);
    }
  }
  hits[q] = hit;
}

}  // namespace rte
