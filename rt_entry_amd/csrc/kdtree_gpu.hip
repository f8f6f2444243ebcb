// kd-tree over the first three coordinates of the rows of a device array,
// built on the GPU.  Same output as the host build (kdtree.cpp in this
// library) and the reference's Kdtree + computeAabbPid (kdtree.cpp:4-195):
// the same leaf point sets, the same tight boxes (as values: where a
// coordinate is +0 in some points and -0 in others, the sign of a zero box
// bound may differ), the same left-first DFS leaf order and the same
// ascending, np-padded pid table.
//
// The rules (kdtree.cpp:30-62, 103-125, 171-185): a node with at most
// max_node points is a leaf; otherwise it splits its tight box's widest axis
// (ties to the lowest axis) at the bisection value of findxM, and values
// below the split go left.  Every output depends only on the set of points
// in a node, never on their order, which the parallel partitions below use.
//
// Large nodes (more than kKdSmall points) are split a level at a time:
//   box       one workgroup per 4096-point chunk of a node, min/max of
//             order-preserving integer keys of the coordinates, atomics per node
//   bisection the split-axis coordinates of the level are gathered once into
//             a contiguous array; every node of the level runs findxM's loop
//             in lockstep: one counting kernel (chunks, one atomic per
//             workgroup) and one per-node update kernel per step; the host
//             stops the steps when no node is left (read back every fourth
//             step)
//   partition slots from both ends of the node's segment, one atomic per
//             side and workgroup
// Small nodes (at most kKdSmall points) are finished by one workgroup each:
// ids, coordinates and a local permutation in LDS, an explicit stack of
// segments, block-wide counts for the bisection.  Left children always come
// before right children in the index array, so the left-first DFS order of
// the leaves is the order of their segment starts.  Each leaf's ids are
// sorted by a bitonic network in LDS to form its pid row.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "ctx.h"

namespace rte {

namespace {

constexpr int kKdSmall = 2048;   // nodes up to this size: one workgroup in LDS
constexpr int kKdChunk = 4096;   // points per workgroup in the level kernels
constexpr int kKdMaxIter = 64;   // findxM ends within about 11 steps

struct LevelNode {
  int64_t l, r;       // segment [l, r) of the index array
  int64_t blk0;       // first workgroup of this node in the level grids
  unsigned key[6];    // ordered keys: min x, y, z, max x, y, z
  int axis, done;
  float bl, br, xm, eps;
  unsigned cnt, nl, nr, pad;
};

struct LeafRec {
  int64_t start;
  int cnt, pad;
  float box[6];
};

__device__ __forceinline__ unsigned fkey(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float funkey(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// findxM's midpoint (kdtree.cpp:103-125): the float sum, halved in double
__device__ __forceinline__ float kd_mid(float l, float r) {
  return (float)((double)(r + l) / 2.0);
}

__device__ __forceinline__ int node_of_block(const LevelNode* nodes, int m, int64_t b) {
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (nodes[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// 6 keys: min of the first three, max of the last three, over the workgroup
__device__ __forceinline__ void block_minmax6(unsigned v[6], unsigned (*red)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = min(v[k], (unsigned)__shfl_xor((int)v[k], o));
#pragma unroll
    for (int k = 3; k < 6; ++k) v[k] = max(v[k], (unsigned)__shfl_xor((int)v[k], o));
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][wave] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = min(min(red[k][0], red[k][1]), min(red[k][2], red[k][3]));
#pragma unroll
  for (int k = 3; k < 6; ++k) v[k] = max(max(red[k][0], red[k][1]), max(red[k][2], red[k][3]));
  __syncthreads();
}

__device__ __forceinline__ unsigned block_sum(unsigned v, unsigned* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += (unsigned)__shfl_xor((int)v, o);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  v = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return v;
}

// widest axis (strictly wider wins, so ties go to the lowest axis) and the
// initial bisection state (kdtree.cpp:30-42, 103-110)
__device__ __forceinline__ void split_setup(const float lo[3], const float hi[3], int& axis,
                                            float& bl, float& br, float& xm, float& eps) {
  float maxw = 0.0f;
  axis = 0;
  for (int a = 0; a < 3; ++a)
    if (hi[a] - lo[a] > maxw) {
      maxw = hi[a] - lo[a];
      axis = a;
    }
  bl = lo[axis];
  br = hi[axis];
  xm = kd_mid(bl, br);
  eps = fmaxf((br - bl) / 1000.0f, 0.00001f);
}

__global__ void kd_iota_kernel(int* idx, int64_t np) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) idx[i] = (int)i;
}

__global__ void kd_level_init_kernel(LevelNode* nodes, int m, unsigned* act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kKdMaxIter) act[i] = 0;
  if (i >= m) return;
  LevelNode& n = nodes[i];
  for (int k = 0; k < 3; ++k) {
    n.key[k] = 0xffffffffu;
    n.key[3 + k] = 0u;
  }
  n.cnt = n.nl = n.nr = 0;
  n.done = 0;
}

__global__ __launch_bounds__(256) void kd_box_kernel(const float* __restrict__ pts, int64_t stride,
                                                     const int* __restrict__ idx,
                                                     LevelNode* nodes, int m) {
  __shared__ unsigned red[6][4];
  const int nd = node_of_block(nodes, m, blockIdx.x);
  const int64_t s0 = nodes[nd].l + ((int64_t)blockIdx.x - nodes[nd].blk0) * kKdChunk;
  const int64_t s1 = min(nodes[nd].r, s0 + kKdChunk);
  unsigned v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (int64_t i = s0 + threadIdx.x; i < s1; i += 256) {
    const float* p = pts + (int64_t)idx[i] * stride;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const unsigned k = fkey(p[a]);
      v[a] = min(v[a], k);
      v[3 + a] = max(v[3 + a], k);
    }
  }
  block_minmax6(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&nodes[nd].key[a], v[a]);
      atomicMax(&nodes[nd].key[3 + a], v[3 + a]);
    }
  }
}

__global__ void kd_split_init_kernel(LevelNode* nodes, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LevelNode& n = nodes[i];
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = funkey(n.key[a]);
    hi[a] = funkey(n.key[3 + a]);
  }
  int axis;
  float bl, br, xm, eps;
  split_setup(lo, hi, axis, bl, br, xm, eps);
  n.axis = axis;
  n.bl = bl;
  n.br = br;
  n.xm = xm;
  n.eps = eps;
  n.done = !(br - bl > eps);
}

// vals[i] = the split-axis coordinate of point idx[i], once per level, so
// the bisection steps and the partition read it contiguously
__global__ __launch_bounds__(256) void kd_gather_kernel(const float* __restrict__ pts,
                                                        int64_t stride,
                                                        const int* __restrict__ idx,
                                                        const LevelNode* nodes, int m,
                                                        float* __restrict__ vals) {
  const int nd = node_of_block(nodes, m, blockIdx.x);
  const int axis = nodes[nd].axis;
  const int64_t s0 = nodes[nd].l + ((int64_t)blockIdx.x - nodes[nd].blk0) * kKdChunk;
  const int64_t s1 = min(nodes[nd].r, s0 + kKdChunk);
  for (int64_t i = s0 + threadIdx.x; i < s1; i += 256)
    vals[i] = pts[(int64_t)idx[i] * stride + axis];
}

__global__ __launch_bounds__(256) void kd_count_kernel(const float* __restrict__ vals,
                                                       LevelNode* nodes, int m) {
  __shared__ unsigned red[4];
  const int nd = node_of_block(nodes, m, blockIdx.x);
  if (nodes[nd].done) return;
  const float mid = kd_mid(nodes[nd].bl, nodes[nd].br);
  const int64_t s0 = nodes[nd].l + ((int64_t)blockIdx.x - nodes[nd].blk0) * kKdChunk;
  const int64_t s1 = min(nodes[nd].r, s0 + kKdChunk);
  unsigned c = 0;
  for (int64_t i = s0 + threadIdx.x; i < s1; i += 256) c += vals[i] <= mid;
  c = block_sum(c, red);
  if (threadIdx.x == 0 && c) atomicAdd(&nodes[nd].cnt, c);
}

// one findxM step per node (kdtree.cpp:111-124); act counts nodes still going
__global__ void kd_update_kernel(LevelNode* nodes, int m, unsigned* act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  LevelNode& n = nodes[i];
  if (n.done) return;
  const float mid = kd_mid(n.bl, n.br);
  const int64_t half = (n.r - n.l) / 2;
  const int64_t cnt = n.cnt;
  n.cnt = 0;
  if (cnt <= half) {
    n.bl = mid;
    n.xm = mid;
    if (cnt == half) n.done = 1;
  } else {
    n.br = mid;
  }
  if (!n.done && !(n.br - n.bl > n.eps)) n.done = 1;
  if (!n.done) atomicAdd(act, 1u);
}

// Points of a chunk below the split go to the front of the node's segment,
// the rest to the back.  Each thread holds its 16 points of the chunk; the
// workgroup counts both sides, takes its slots with one atomic per side and
// node, and the waves place their points from per-wave bases.
__global__ __launch_bounds__(256) void kd_partition_kernel(const float* __restrict__ vals,
                                                           const int* __restrict__ idx,
                                                           int* __restrict__ idx2,
                                                           LevelNode* nodes, int m) {
  constexpr int PER = kKdChunk / 256;
  __shared__ unsigned cnt_w[2][4];
  __shared__ unsigned base_w[2][4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nd = node_of_block(nodes, m, blockIdx.x);
  const int64_t l = nodes[nd].l, r = nodes[nd].r;
  const float xm = nodes[nd].xm;
  const int64_t s0 = l + ((int64_t)blockIdx.x - nodes[nd].blk0) * kKdChunk;
  const int64_t s1 = min(r, s0 + kKdChunk);
  int id[PER];
  unsigned lmask = 0, vmask = 0;  // bit k: element k is valid / goes left
  unsigned nl = 0, nr = 0;        // this wave's counts (uniform)
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t i = s0 + 256 * k + t;
    const bool valid = i < s1;
    id[k] = valid ? idx[i] : 0;
    const bool left = valid && vals[i] < xm;
    vmask |= (unsigned)valid << k;
    lmask |= (unsigned)left << k;
    nl += (unsigned)__popcll(__ballot(left));
    nr += (unsigned)__popcll(__ballot(valid && !left));
  }
  if (lane == 0) {
    cnt_w[0][wave] = nl;
    cnt_w[1][wave] = nr;
  }
  __syncthreads();
  if (t == 0) {
    const unsigned tl = cnt_w[0][0] + cnt_w[0][1] + cnt_w[0][2] + cnt_w[0][3];
    const unsigned tr = cnt_w[1][0] + cnt_w[1][1] + cnt_w[1][2] + cnt_w[1][3];
    unsigned bl = tl ? atomicAdd(&nodes[nd].nl, tl) : 0u;
    unsigned br = tr ? atomicAdd(&nodes[nd].nr, tr) : 0u;
    for (int w = 0; w < 4; ++w) {
      base_w[0][w] = bl;
      base_w[1][w] = br;
      bl += cnt_w[0][w];
      br += cnt_w[1][w];
    }
  }
  __syncthreads();
  unsigned ol = base_w[0][wave], orr = base_w[1][wave];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const bool valid = (vmask >> k) & 1u, left = (lmask >> k) & 1u;
    const unsigned long long ml = __ballot(left), mr = __ballot(valid && !left);
    if (left) idx2[l + ol + lanes_below(ml)] = id[k];
    if (valid && !left) idx2[r - 1 - (int64_t)(orr + lanes_below(mr))] = id[k];
    ol += (unsigned)__popcll(ml);
    orr += (unsigned)__popcll(mr);
  }
}

__global__ __launch_bounds__(256) void kd_copy_kernel(const int* __restrict__ idx2,
                                                      int* __restrict__ idx,
                                                      const LevelNode* nodes, int m) {
  const int nd = node_of_block(nodes, m, blockIdx.x);
  const int64_t s0 = nodes[nd].l + ((int64_t)blockIdx.x - nodes[nd].blk0) * kKdChunk;
  const int64_t s1 = min(nodes[nd].r, s0 + kKdChunk);
  for (int64_t i = s0 + threadIdx.x; i < s1; i += 256) idx[i] = idx2[i];
}

// err bits: 1 = split failed, 2 = findxM did not end, 4 = leaf capacity
__global__ __launch_bounds__(256) void kd_small_kernel(const float* __restrict__ pts,
                                                       int64_t stride, int* __restrict__ idx,
                                                       const int64_t* __restrict__ segs,
                                                       int max_node, LeafRec* __restrict__ leaves,
                                                       unsigned* leaf_ctr, int max_leaves,
                                                       unsigned* err) {
  __shared__ int ids[kKdSmall];
  __shared__ float cs[3][kKdSmall];
  __shared__ uint16_t perm[2][kKdSmall];
  __shared__ unsigned stack[kKdSmall];  // a << 16 | b, local offsets
  __shared__ unsigned red[6][4];
  __shared__ unsigned cnts[2];
  const int t = threadIdx.x;
  const int64_t l = segs[2 * blockIdx.x];
  const int n = (int)(segs[2 * blockIdx.x + 1] - l);
  for (int j = t; j < n; j += 256) {
    const int id = idx[l + j];
    ids[j] = id;
    const float* p = pts + (int64_t)id * stride;
    cs[0][j] = p[0];
    cs[1][j] = p[1];
    cs[2][j] = p[2];
    perm[0][j] = (uint16_t)j;
  }
  if (t == 0) stack[0] = (unsigned)n;
  int top = 1;
  __syncthreads();
  while (top > 0) {
    const unsigned e = stack[--top];
    const int a = (int)(e >> 16), b = (int)(e & 0xffffu);
    const int cnt = b - a;
    unsigned v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (int j = a + t; j < b; j += 256) {
      const int p = perm[0][j];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const unsigned k = fkey(cs[ax][p]);
        v[ax] = min(v[ax], k);
        v[3 + ax] = max(v[3 + ax], k);
      }
    }
    block_minmax6(v, red);
    float lo[3], hi[3];
    for (int ax = 0; ax < 3; ++ax) {
      lo[ax] = funkey(v[ax]);
      hi[ax] = funkey(v[3 + ax]);
    }
    if (cnt <= max_node) {  // leaf (kdtree.cpp:171-174)
      if (t == 0) {
        const unsigned s = atomicAdd(leaf_ctr, 1u);
        if (s < (unsigned)max_leaves) {
          LeafRec& rec = leaves[s];
          rec.start = l + a;
          rec.cnt = cnt;
          rec.pad = 0;
          for (int ax = 0; ax < 3; ++ax) {
            rec.box[ax] = lo[ax];
            rec.box[3 + ax] = hi[ax];
          }
        } else {
          atomicOr(err, 4u);
        }
      }
      for (int j = a + t; j < b; j += 256) idx[l + j] = ids[perm[0][j]];
      __syncthreads();
      continue;
    }
    int axis;
    float bl, br, xm, eps;
    split_setup(lo, hi, axis, bl, br, xm, eps);
    const unsigned half = (unsigned)cnt / 2;
    int steps = 0;
    while (br - bl > eps) {  // findxM (kdtree.cpp:111-124), block-uniform
      if (++steps > kKdMaxIter) break;
      const float mid = kd_mid(bl, br);
      unsigned c = 0;
      for (int j = a + t; j < b; j += 256) c += cs[axis][perm[0][j]] <= mid;
      c = block_sum(c, &red[0][0]);
      if (c <= half) {
        bl = mid;
        xm = mid;
        if (c == half) break;
      } else {
        br = mid;
      }
    }
    if (steps > kKdMaxIter) {
      if (t == 0) atomicOr(err, 2u);
      return;
    }
    if (t < 2) cnts[t] = 0;
    __syncthreads();
    for (int base = a; base < b; base += 256) {
      const int j = base + t;
      const bool valid = j < b;
      const int p = valid ? perm[0][j] : 0;
      const bool left = valid && cs[axis][p] < xm;
      const bool right = valid && !left;
      const unsigned long long ml = __ballot(left), mr = __ballot(right);
      unsigned lb = 0, rb = 0;
      if ((t & 63) == 0) {
        if (ml) lb = atomicAdd(&cnts[0], (unsigned)__popcll(ml));
        if (mr) rb = atomicAdd(&cnts[1], (unsigned)__popcll(mr));
      }
      lb = (unsigned)__builtin_amdgcn_readfirstlane((int)lb);
      rb = (unsigned)__builtin_amdgcn_readfirstlane((int)rb);
      if (left) perm[1][a + lb + lanes_below(ml)] = (uint16_t)p;
      if (right) perm[1][b - 1 - (int)(rb + lanes_below(mr))] = (uint16_t)p;
    }
    __syncthreads();
    const int nl = (int)cnts[0];
    for (int j = a + t; j < b; j += 256) perm[0][j] = perm[1][j];
    if (nl == 0 || nl == cnt) {  // kdtree.cpp:182-185
      if (t == 0) atomicOr(err, 1u);
      return;
    }
    if (t == 0) {
      stack[top] = ((unsigned)(a + nl) << 16) | (unsigned)b;
      stack[top + 1] = ((unsigned)a << 16) | (unsigned)(a + nl);
    }
    top += 2;
    __syncthreads();
  }
}

// pid row of leaf i: its ids ascending, padded with np (kdtree.cpp:91-101)
__global__ __launch_bounds__(256) void kd_pids_kernel(const int* __restrict__ idx,
                                                      const int64_t* __restrict__ start,
                                                      const int* __restrict__ count, int max_node,
                                                      int P, int64_t np,
                                                      int32_t* __restrict__ pids) {
  __shared__ int s[kKdSmall];
  const int leaf = blockIdx.x;
  const int64_t st = start[leaf];
  const int cnt = count[leaf];
  for (int j = threadIdx.x; j < P; j += 256) s[j] = j < cnt ? idx[st + j] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < P; i += 256) {
        const int ixj = i ^ jj;
        if (ixj > i) {
          const int x = s[i], y = s[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            s[i] = y;
            s[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int j = threadIdx.x; j < max_node; j += 256)
    pids[(int64_t)leaf * max_node + j] = j < cnt ? s[j] : (int32_t)np;
}

unsigned blocks_1d(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

}  // namespace rte

extern "C" {

int rte_build_kdtree_gpu(rte_ctx* c, const float* pts, int64_t np, int64_t stride, int max_node,
                         int max_leaves, float* aabbs_out, int32_t* pids_out, int* n_leaf_out,
                         int where) {
  using namespace rte;
  if (!c) return RTE_EINVAL;
  if (!pts || !aabbs_out || !pids_out || !n_leaf_out || np <= 0 || max_node <= 0 ||
      max_leaves <= 0 || stride < 3)
    return rte_fail(c, RTE_EINVAL, "bad kd-tree arguments");
  *n_leaf_out = 0;
  if (np > INT_MAX) return rte_fail(c, RTE_EINVAL, "kd-tree: np %lld exceeds int32 ids", (long long)np);
  if (max_node > kKdSmall)
    return rte_fail(c, RTE_EUNSUP, "GPU kd-tree: max_node %d > %d (use rte_build_kdtree)",
                    max_node, kKdSmall);
  if (hipSetDevice(c->device) != hipSuccess) return rte_fail(c, RTE_EHIP, "hipSetDevice");
  hipStream_t st = c->stream;
  auto ck = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? RTE_OK : rte_fail(c, RTE_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
#define KCK(expr, what) \
  do {                                  \
    if (int r_ = ck((expr), what)) return r_; \
  } while (0)
  DevBuf pin, idxb, idx2b, valsb, nodesb, actb, segb, leafb, ctrb, ordb, pidb;
  const float* dp = pts;
  int64_t ds = stride;
  if (where == RTE_HOST) {  // only the three coordinates travel
    std::vector<float> p3((size_t)np * 3);
    for (int64_t i = 0; i < np; ++i)
      for (int a = 0; a < 3; ++a) p3[(size_t)i * 3 + a] = pts[i * stride + a];
    KCK(pin.reserve(p3.size() * sizeof(float)), "hipMalloc");
    KCK(hipMemcpyAsync(pin.p, p3.data(), p3.size() * sizeof(float), hipMemcpyHostToDevice, st),
        "copy");
    KCK(hipStreamSynchronize(st), "sync");
    dp = pin.as<float>();
    ds = 3;
  }
  KCK(idxb.reserve((size_t)np * 4), "hipMalloc");
  KCK(idx2b.reserve((size_t)np * 4), "hipMalloc");
  KCK(valsb.reserve((size_t)np * 4), "hipMalloc");
  float* vals = valsb.as<float>();
  KCK(actb.reserve(kKdMaxIter * sizeof(unsigned)), "hipMalloc");
  KCK(ctrb.reserve(2 * sizeof(unsigned)), "hipMalloc");
  int* idx = idxb.as<int>();
  int* idx2 = idx2b.as<int>();
  unsigned* act = actb.as<unsigned>();
  hipLaunchKernelGGL(kd_iota_kernel, dim3(blocks_1d(np, 256)), dim3(256), 0, st, idx, np);

  std::vector<std::pair<int64_t, int64_t>> level, small;
  (np <= kKdSmall ? small : level).push_back({0, np});
  std::vector<LevelNode> hn;
  while (!level.empty()) {
    const int m = (int)level.size();
    hn.assign(m, LevelNode{});
    int64_t blocks = 0;
    for (int i = 0; i < m; ++i) {
      hn[i].l = level[i].first;
      hn[i].r = level[i].second;
      hn[i].blk0 = blocks;
      blocks += (hn[i].r - hn[i].l + kKdChunk - 1) / kKdChunk;
    }
    if (blocks > INT_MAX) return rte_fail(c, RTE_EINVAL, "kd-tree level too large");
    KCK(nodesb.reserve((size_t)m * sizeof(LevelNode)), "hipMalloc");
    LevelNode* nodes = nodesb.as<LevelNode>();
    KCK(hipMemcpyAsync(nodes, hn.data(), (size_t)m * sizeof(LevelNode), hipMemcpyHostToDevice, st),
        "copy");
    const unsigned gm = blocks_1d(std::max(m, kKdMaxIter), 256);
    hipLaunchKernelGGL(kd_level_init_kernel, dim3(gm), dim3(256), 0, st, nodes, m, act);
    hipLaunchKernelGGL(kd_box_kernel, dim3((unsigned)blocks), dim3(256), 0, st, dp, ds, idx, nodes,
                       m);
    hipLaunchKernelGGL(kd_split_init_kernel, dim3(gm), dim3(256), 0, st, nodes, m);
    hipLaunchKernelGGL(kd_gather_kernel, dim3((unsigned)blocks), dim3(256), 0, st, dp, ds, idx,
                       nodes, m, vals);
    bool converged = false;
    for (int it = 0; it < kKdMaxIter; ++it) {
      hipLaunchKernelGGL(kd_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, vals, nodes,
                         m);
      hipLaunchKernelGGL(kd_update_kernel, dim3(gm), dim3(256), 0, st, nodes, m, act + it);
      if ((it & 3) == 3 || it == kKdMaxIter - 1) {
        unsigned left = 0;
        KCK(hipMemcpyAsync(&left, act + it, sizeof(unsigned), hipMemcpyDeviceToHost, st), "copy");
        KCK(hipStreamSynchronize(st), "sync");
        if (left == 0) {
          converged = true;
          break;
        }
      }
    }
    if (!converged) return rte_fail(c, RTE_EBUILD, "kd-tree split search did not end");
    hipLaunchKernelGGL(kd_partition_kernel, dim3((unsigned)blocks), dim3(256), 0, st, vals, idx,
                       idx2, nodes, m);
    hipLaunchKernelGGL(kd_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, idx2, idx, nodes,
                       m);
    KCK(hipGetLastError(), "kd-tree level kernels");
    KCK(hipMemcpyAsync(hn.data(), nodes, (size_t)m * sizeof(LevelNode), hipMemcpyDeviceToHost, st),
        "copy");
    KCK(hipStreamSynchronize(st), "sync");
    std::vector<std::pair<int64_t, int64_t>> next;
    for (int i = 0; i < m; ++i) {
      const int64_t l = hn[i].l, r = hn[i].r, mid = l + hn[i].nl;
      if (mid == l || mid == r) return rte_fail(c, RTE_EBUILD, "kd-tree split failed");
      for (auto ch : {std::make_pair(l, mid), std::make_pair(mid, r)})
        (ch.second - ch.first <= kKdSmall ? small : next).push_back(ch);
    }
    level.swap(next);
  }

  // small nodes and leaves
  const int ns = (int)small.size();
  KCK(segb.reserve((size_t)ns * 2 * sizeof(int64_t)), "hipMalloc");
  KCK(leafb.reserve((size_t)max_leaves * sizeof(LeafRec)), "hipMalloc");
  std::vector<int64_t> segs((size_t)ns * 2);
  for (int i = 0; i < ns; ++i) {
    segs[2 * i] = small[i].first;
    segs[2 * i + 1] = small[i].second;
  }
  KCK(hipMemcpyAsync(segb.p, segs.data(), segs.size() * sizeof(int64_t), hipMemcpyHostToDevice, st),
      "copy");
  KCK(hipMemsetAsync(ctrb.p, 0, 2 * sizeof(unsigned), st), "memset");
  unsigned* leaf_ctr = ctrb.as<unsigned>();
  hipLaunchKernelGGL(kd_small_kernel, dim3(ns), dim3(256), 0, st, dp, ds, idx,
                     segb.as<int64_t>(), max_node, leafb.as<LeafRec>(), leaf_ctr, max_leaves,
                     leaf_ctr + 1);
  KCK(hipGetLastError(), "kd_small_kernel");
  unsigned hc[2];
  KCK(hipMemcpyAsync(hc, ctrb.p, sizeof(hc), hipMemcpyDeviceToHost, st), "copy");
  KCK(hipStreamSynchronize(st), "sync");
  if (hc[1] & 1u) return rte_fail(c, RTE_EBUILD, "kd-tree split failed");
  if (hc[1] & 2u) return rte_fail(c, RTE_EBUILD, "kd-tree split search did not end");
  if ((hc[1] & 4u) || hc[0] > (unsigned)max_leaves)
    return rte_fail(c, RTE_EINVAL, "kd-tree: more than max_leaves=%d leaves", max_leaves);
  const int nl = (int)hc[0];
  std::vector<LeafRec> recs(nl);
  KCK(hipMemcpyAsync(recs.data(), leafb.p, (size_t)nl * sizeof(LeafRec), hipMemcpyDeviceToHost,
                     st),
      "copy");
  KCK(hipStreamSynchronize(st), "sync");
  std::sort(recs.begin(), recs.end(),
            [](const LeafRec& x, const LeafRec& y) { return x.start < y.start; });
  std::vector<float> boxes((size_t)nl * 6);
  std::vector<int64_t> starts(nl);
  std::vector<int> counts(nl);
  for (int i = 0; i < nl; ++i) {
    std::copy(recs[i].box, recs[i].box + 6, boxes.begin() + (size_t)i * 6);
    starts[i] = recs[i].start;
    counts[i] = recs[i].cnt;
  }
  KCK(ordb.reserve((size_t)nl * (sizeof(int64_t) + sizeof(int))), "hipMalloc");
  int64_t* dstart = ordb.as<int64_t>();
  int* dcount = reinterpret_cast<int*>(dstart + nl);
  KCK(hipMemcpyAsync(dstart, starts.data(), (size_t)nl * sizeof(int64_t), hipMemcpyHostToDevice,
                     st),
      "copy");
  KCK(hipMemcpyAsync(dcount, counts.data(), (size_t)nl * sizeof(int), hipMemcpyHostToDevice, st),
      "copy");
  int P = 1;
  while (P < max_node) P <<= 1;
  int32_t* dpids = pids_out;
  if (where == RTE_HOST) {
    KCK(pidb.reserve((size_t)nl * max_node * sizeof(int32_t)), "hipMalloc");
    dpids = pidb.as<int32_t>();
  }
  hipLaunchKernelGGL(kd_pids_kernel, dim3(nl), dim3(256), 0, st, idx, dstart, dcount, max_node, P,
                     np, dpids);
  KCK(hipGetLastError(), "kd_pids_kernel");
  if (where == RTE_HOST) {
    KCK(hipMemcpyAsync(pids_out, dpids, (size_t)nl * max_node * sizeof(int32_t),
                       hipMemcpyDeviceToHost, st),
        "copy");
    std::copy(boxes.begin(), boxes.end(), aabbs_out);
  } else {
    KCK(hipMemcpyAsync(aabbs_out, boxes.data(), boxes.size() * sizeof(float),
                       hipMemcpyHostToDevice, st),
        "copy");
  }
  KCK(hipStreamSynchronize(st), "sync");
#undef KCK
  *n_leaf_out = nl;
  return RTE_OK;
}

}  // extern "C"
