// Full-precision rerank of search candidates.
//
// Replaces Graph::parallel_reorder (graph_helper.cu:146-213, run on the host
// with OpenMP after the search when REORDER is defined, graph.cu:326-336):
// the search runs on the first DIM PCA dimensions and returns its whole beam
// (graph_search.cuh:1955-1958); every candidate's distance is then recomputed
// in the original d-dimensional space and the k best are kept.
//
// One wave per query, one candidate per lane (up to 8 per lane).  A lane
// computes its candidate's distance exactly as the reference's AVX loop:
// eight partial sums over dims l, l + 8, l + 16, ... (product, then add),
// added in order 0..7 at the end, so the values are the reference's bits.
// The k best by (dist, id) are then selected by k rounds of a wave-wide
// minimum.  Ids outside [0, np) (unfilled beam slots) get +inf.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "rerank.h"

namespace rte {

__device__ __forceinline__ bool key_less(float da, int ia, float db, int ib) {
  return da < db || (da == db && ia < ib);
}

template <int R>
__global__ __launch_bounds__(256) void rerank_kernel(const float* __restrict__ base, int64_t np,
                                                     int64_t stride, int d,
                                                     const float* __restrict__ Q, int64_t nq,
                                                     int64_t q_stride,
                                                     const int32_t* __restrict__ cand, int n_cand,
                                                     int k, int32_t* __restrict__ ids_out,
                                                     float* __restrict__ dists_out) {
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (q >= nq) return;  // whole waves; no workgroup barrier below
  const float* qv = Q + q * q_stride;
  float dist[R];
  int id[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = lane + 64 * r;
    id[r] = 0x7fffffff;
    dist[r] = INFINITY;
    if (c < n_cand) {
      const int32_t x = cand[q * n_cand + c];
      id[r] = x;
      if (x >= 0 && x < np) {
        const float* pv = base + (int64_t)x * stride;
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int j = 0; j < d; j += 8) {
          const float4 p0 = *reinterpret_cast<const float4*>(pv + j);
          const float4 p1 = *reinterpret_cast<const float4*>(pv + j + 4);
          const float4 q0 = *reinterpret_cast<const float4*>(qv + j);
          const float4 q1 = *reinterpret_cast<const float4*>(qv + j + 4);
          const float e[8] = {p0.x - q0.x, p0.y - q0.y, p0.z - q0.z, p0.w - q0.w,
                              p1.x - q1.x, p1.y - q1.y, p1.z - q1.z, p1.w - q1.w};
#pragma unroll
          for (int l = 0; l < 8; ++l) acc[l] = acc[l] + e[l] * e[l];
        }
        float s = 0.f;
#pragma unroll
        for (int l = 0; l < 8; ++l) s = s + acc[l];
        dist[r] = s;
      }
    }
  }
  for (int out = 0; out < k; ++out) {
    float bd = INFINITY;
    int bi = 0x7fffffff, br = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (key_less(dist[r], id[r], bd, bi)) {
        bd = dist[r];
        bi = id[r];
        br = r;
      }
    float md = bd;
    int mi = bi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float od = __shfl_xor(md, o, 64);
      const int oi = __shfl_xor(mi, o, 64);
      if (key_less(od, oi, md, mi)) {
        md = od;
        mi = oi;
      }
    }
    // the lowest lane holding the minimum key gives up that slot
    const uint64_t mine = __ballot(bd == md && bi == mi);
    if (lane == __ffsll((unsigned long long)mine) - 1) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r == br) {
          dist[r] = INFINITY;
          id[r] = 0x7fffffff;
        }
      ids_out[q * k + out] = mi;
      if (dists_out) dists_out[q * k + out] = md;
    }
  }
}

hipError_t launch_rerank(const float* base, int64_t np, int64_t stride, int d, const float* Q,
                         int64_t nq, int64_t q_stride, const int32_t* cand, int n_cand, int k,
                         int32_t* ids_out, float* dists_out, hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nq + 3) / 4));
  const int r = (n_cand + 63) / 64;
#define RR(N)                                                                                    \
  hipLaunchKernelGGL(rerank_kernel<N>, grid, dim3(256), 0, st, base, np, stride, d, Q, nq,      \
                     q_stride, cand, n_cand, k, ids_out, dists_out)
  if (r <= 1) RR(1);
  else if (r <= 2) RR(2);
  else if (r <= 4) RR(4);
  else RR(8);
#undef RR
  return hipGetLastError();
}

}  // namespace rte
