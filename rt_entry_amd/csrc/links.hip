// Entry-space links: the last r slots of every graph row become points drawn
// from the kd-tree leaves around the row's own leaf.
//
// Leaves are numbered in kd-tree DFS order, so leaves [l - w, l + w] are a
// neighbourhood of leaf l in the 3-D PCA space that entry selection works in.
// On data made of many well-separated clusters, a kNN graph has no edges
// between clusters, and a query whose entry leaf holds no point of its own
// cluster among the beam's seeds never reaches it.  These links let a search
// that starts in the wrong cluster reach the points that share the entry
// region with the query.
//
// Draws are deterministic: candidate t for slot j of node u is
// splitmix64(seed + ((u << 16) | (j << 8) | t) * 0x9E3779B97F4A7C15) modulo
// the number of points in the window, counted leaf by leaf in the pid table's
// order; a candidate equal to u, to one of the row's kept entries or to an
// earlier link is redrawn (t + 1), and after kLinkTries draws the slot is
// padding (np).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "links.h"

namespace rte {

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ull;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z;
}

// one thread per pid-table slot: the leaf of every point and every leaf's
// size (its rows are ascending ids padded with np at the end)
__global__ void leaf_index_kernel(const int32_t* __restrict__ pids, int n_leaf, int cap, int64_t np,
                                  int32_t* __restrict__ leaf_of, int32_t* __restrict__ size) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_leaf * cap) return;
  const int l = (int)(t / cap), i = (int)(t % cap);
  const int32_t p = pids[t];
  if (p < 0 || p >= np) return;
  leaf_of[p] = l;
  if (i == cap - 1 || pids[t + 1] >= np) size[l] = i + 1;
}

__global__ void entry_links_kernel(int32_t* __restrict__ graph, int64_t np, int degree, int r,
                                   int window, uint64_t seed, const int32_t* __restrict__ pids,
                                   int n_leaf, int cap, const int32_t* __restrict__ leaf_of,
                                   const int32_t* __restrict__ size) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= np) return;
  int32_t* row = graph + u * degree;
  const int keep = degree - r;
  const int l = leaf_of[u];
  if (l < 0) {  // a point in no leaf keeps its row (padded links)
    for (int j = 0; j < r; ++j) row[keep + j] = (int32_t)np;
    return;
  }
  const int lo = l - window < 0 ? 0 : l - window;
  const int hi = l + window > n_leaf - 1 ? n_leaf - 1 : l + window;
  uint64_t total = 0;
  for (int x = lo; x <= hi; ++x) total += (uint64_t)size[x];
  int32_t links[kMaxLinks];
  for (int j = 0; j < r; ++j) {
    int32_t pick = (int32_t)np;
    for (int t = 0; t < kLinkTries; ++t) {
      const uint64_t key = ((uint64_t)u << 16) | ((uint64_t)j << 8) | (uint64_t)t;
      uint64_t idx = splitmix64(seed + key * 0x9E3779B97F4A7C15ull) % total;
      int x = lo;
      while (idx >= (uint64_t)size[x]) {
        idx -= (uint64_t)size[x];
        ++x;
      }
      const int32_t c = pids[(int64_t)x * cap + (int64_t)idx];
      bool dup = c == (int32_t)u;
      for (int i = 0; i < keep && !dup; ++i) dup = row[i] == c;
      for (int i = 0; i < j && !dup; ++i) dup = links[i] == c;
      if (!dup) {
        pick = c;
        break;
      }
    }
    links[j] = pick;
  }
  for (int j = 0; j < r; ++j) row[keep + j] = links[j];
}

hipError_t launch_entry_links(int32_t* graph, int64_t np, int degree, int r, int window,
                              uint64_t seed, const int32_t* pids, int n_leaf, int cap,
                              int32_t* leaf_of, int32_t* size, hipStream_t st) {
  hipError_t e = hipMemsetAsync(leaf_of, 0xff, (size_t)np * sizeof(int32_t), st);
  if (e == hipSuccess) e = hipMemsetAsync(size, 0, (size_t)n_leaf * sizeof(int32_t), st);
  if (e != hipSuccess) return e;
  const int64_t slots = (int64_t)n_leaf * cap;
  hipLaunchKernelGGL(leaf_index_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, st,
                     pids, n_leaf, cap, np, leaf_of, size);
  hipLaunchKernelGGL(entry_links_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st,
                     graph, np, degree, r, window, seed, pids, n_leaf, cap, leaf_of, size);
  return hipGetLastError();
}

}  // namespace rte
