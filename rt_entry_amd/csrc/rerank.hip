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

// LDS staging of the candidate rows: 64 dims of the 64 candidates of a
// round at a time, rows padded by 4 floats (a lane's 16-byte reads of its
// own row then spread over the banks), followed by the same 64 query dims
constexpr int kRrChunk = 64;
constexpr int kRrLd = kRrChunk + 4;
constexpr int kRrWaveFloats = 64 * kRrLd + kRrChunk;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One lane's candidate distance from the staged rows: eight partial sums
// over dims l, l + 8, ... (product, then add) continued chunk after chunk in
// ascending dimension order, exactly as the unstaged loop.
template <int R>
__device__ __forceinline__ void staged_dists(const float* __restrict__ base, int64_t np,
                                             int64_t stride, int d, const float* __restrict__ qv,
                                             int n_cand, int lane, float* rows,
                                             const int (&id)[R], float (&dist)[R]) {
  float* qs = rows + 64 * kRrLd;
  const int seg = lane & 15, sub = lane >> 4;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    // source rows of this round: lane 16 t' + seg loads row 4 t + sub
    int64_t src[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int row = 4 * t + sub;
      const int x = __shfl(id[r], row, 64);
      src[t] = (row + 64 * r < n_cand && x >= 0 && x < np) ? (int64_t)x * stride : -1;
    }
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < d; j0 += kRrChunk) {
      float4 v[16];
#pragma unroll
      for (int t = 0; t < 16; ++t)
        v[t] = src[t] >= 0 ? *reinterpret_cast<const float4*>(base + src[t] + j0 + 4 * seg)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 qq = *reinterpret_cast<const float4*>(qv + j0 + 4 * seg);
      wave_lds_sync();  // the previous chunk's reads are done
#pragma unroll
      for (int t = 0; t < 16; ++t)
        *reinterpret_cast<float4*>(rows + (4 * t + sub) * kRrLd + 4 * seg) = v[t];
      if (sub == 0) *reinterpret_cast<float4*>(qs + 4 * seg) = qq;
      wave_lds_sync();
      const float* my = rows + lane * kRrLd;
#pragma unroll
      for (int h = 0; h < kRrChunk / 8; ++h) {
        const float4 p0 = *reinterpret_cast<const float4*>(my + 8 * h);
        const float4 p1 = *reinterpret_cast<const float4*>(my + 8 * h + 4);
        const float4 q0 = *reinterpret_cast<const float4*>(qs + 8 * h);
        const float4 q1 = *reinterpret_cast<const float4*>(qs + 8 * h + 4);
        const float e[8] = {p0.x - q0.x, p0.y - q0.y, p0.z - q0.z, p0.w - q0.w,
                            p1.x - q1.x, p1.y - q1.y, p1.z - q1.z, p1.w - q1.w};
#pragma unroll
        for (int l = 0; l < 8; ++l) acc[l] = acc[l] + e[l] * e[l];
      }
    }
    const int c = lane + 64 * r;
    if (c < n_cand && id[r] >= 0 && id[r] < np) {
      float sum = 0.f;
#pragma unroll
      for (int l = 0; l < 8; ++l) sum = sum + acc[l];
      dist[r] = sum;
    }
  }
}

template <int R, bool STAGED>
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
    id[r] = c < n_cand ? cand[q * n_cand + c] : 0x7fffffff;
    dist[r] = INFINITY;
  }
  if constexpr (STAGED) {
    extern __shared__ float rr_lds[];
    staged_dists<R>(base, np, stride, d, qv, n_cand, lane,
                    rr_lds + (threadIdx.x >> 6) * kRrWaveFloats, id, dist);
  } else {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = lane + 64 * r;
    if (c < n_cand) {
      const int32_t x = id[r];
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
  const int r = (n_cand + 63) / 64;
  // rows in whole 64-dim chunks: coalesced 256-byte row segments staged
  // through LDS (two waves per workgroup, 35 KB); otherwise a lane streams
  // its own row
  const bool staged = d % kRrChunk == 0;
  const int wpb = staged ? 2 : 4;
  const dim3 grid((unsigned)((nq + wpb - 1) / wpb));
  const size_t lds = staged ? (size_t)wpb * kRrWaveFloats * sizeof(float) : 0;
#define RR(N)                                                                              \
  do {                                                                                     \
    if (staged)                                                                            \
      hipLaunchKernelGGL((rerank_kernel<N, true>), grid, dim3(64 * wpb), lds, st, base, np, \
                         stride, d, Q, nq, q_stride, cand, n_cand, k, ids_out, dists_out);  \
    else                                                                                   \
      hipLaunchKernelGGL((rerank_kernel<N, false>), grid, dim3(64 * wpb), 0, st, base, np,  \
                         stride, d, Q, nq, q_stride, cand, n_cand, k, ids_out, dists_out);  \
  } while (0)
  if (r <= 1) RR(1);
  else if (r <= 2) RR(2);
  else if (r <= 4) RR(4);
  else RR(8);
#undef RR
  return hipGetLastError();
}

}  // namespace rte
