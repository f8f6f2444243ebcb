// Entry-leaf selection: stackless traversal of a box hierarchy over the leaf
// AABBs, one lane per query.
//
// Replaces the OptiX ray cast (raytracing.cpp:283-334, devicePrograms.cu:29-96):
// the reference casts a zero-length ray from (pq0,pq1,pq2) into a BVH over
// the expanded leaf boxes and keeps the first box reported.  CDNA4 has no
// ray-tracing units; the leaves are already spatially ordered (left-first
// DFS of the kd-tree, kdtree.cpp:193-194), so a binary hierarchy over leaf
// index ranges, laid out in preorder with skip links, visits leaves in
// ascending id order and the first containing leaf is the smallest such id.
// A miss yields leaf 0 (entry.cu:88-89 fills hits with 0).
//
// Memory layout: the nodes are stored breadth-first with first-child and
// skip links (the walk order stays preorder), so the top levels are a prefix
// of the array; each workgroup copies that prefix (kTopNodes nodes, 64 KB)
// into LDS and only the deepest levels are read from global memory.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "entry.h"
#include "entry_impl.h"

namespace rte {

__global__ __launch_bounds__(256) void select_entry_kernel(const BoxNode* __restrict__ nodes,
                                                           int n_nodes,
                                                           const float* __restrict__ pq,
                                                           int64_t nq, int stride,
                                                           int32_t* __restrict__ hits) {
  __shared__ BoxNode top[kTopNodes];
  select_entry_block(nodes, n_nodes, top, pq, nq, stride, hits, (int64_t)blockIdx.x * 256);
}

hipError_t launch_select_entry(const BoxNode* nodes, int n_nodes, const float* pq, int64_t nq,
                               int stride, int32_t* hits, hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  const int64_t grid = (nq + 255) / 256;
  hipLaunchKernelGGL(select_entry_kernel, dim3((unsigned)grid), dim3(256), 0, st, nodes, n_nodes,
                     pq, nq, stride, hits);
  return hipGetLastError();
}

// Preorder hierarchy over leaves [lo, hi): node box = union of leaf boxes.
static int build_rec(const float* aabbs, int lo, int hi, std::vector<BoxNode>& out) {
  const int me = (int)out.size();
  out.push_back(BoxNode{});
  if (hi - lo == 1) {
    BoxNode& n = out[me];
    for (int a = 0; a < 3; ++a) {
      n.lo[a] = aabbs[(size_t)lo * 6 + a];
      n.hi[a] = aabbs[(size_t)lo * 6 + 3 + a];
    }
    n.link = lo;
  } else {
    const int mid = lo + (hi - lo) / 2;
    const int l = build_rec(aabbs, lo, mid, out);
    const int r = build_rec(aabbs, mid, hi, out);
    BoxNode& n = out[me];
    for (int a = 0; a < 3; ++a) {
      n.lo[a] = fminf(out[l].lo[a], out[r].lo[a]);
      n.hi[a] = fmaxf(out[l].hi[a], out[r].hi[a]);
    }
    n.link = -1;
  }
  out[me].skip = (int)out.size();
  return me;
}

void build_hierarchy(const float* aabbs, int n_leaf, std::vector<BoxNode>& out) {
  out.clear();
  if (n_leaf <= 0) return;
  std::vector<BoxNode> pre;  // preorder: first child = index + 1, skip links
  pre.reserve((size_t)2 * n_leaf);
  build_rec(aabbs, 0, n_leaf, pre);
  const int n = (int)pre.size();
  // breadth-first relabelling; children of preorder node p: p + 1 and the
  // skip of p + 1
  std::vector<int> bfs;  // bfs position -> preorder index
  std::vector<int> pos(n + 1, n);  // preorder index -> bfs position (n -> n)
  bfs.reserve(n);
  bfs.push_back(0);
  for (size_t h = 0; h < bfs.size(); ++h) {
    const int p = bfs[h];
    pos[p] = (int)h;
    if (pre[p].link < 0) {
      bfs.push_back(p + 1);
      bfs.push_back(pre[p + 1].skip);
    }
  }
  out.resize(n);
  for (int h = 0; h < n; ++h) {
    const BoxNode& a = pre[bfs[h]];
    BoxNode b = a;
    b.skip = pos[a.skip];
    b.link = a.link >= 0 ? a.link : -(pos[bfs[h] + 1]) - 1;
    out[h] = b;
  }
}

}  // namespace rte
