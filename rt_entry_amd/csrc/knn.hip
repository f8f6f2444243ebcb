// Exact kNN of a base against itself on FP32 MFMA, for building the search
// graph.  The reference has no graph builder (graph.cu:117-121 leaves a TODO;
// its .nsw files come from an external tool, flags.cpp:14).
//
// Distances are |q|^2 + |x|^2 - 2 q.x with q.x from v_mfma_f32_16x16x4_f32
// tiles.  For integer-valued data with |x|^2 < 2^24 (the 1M x 128 workload:
// values 0..255, d = 128) every term is an exact FP32 integer, so the lists
// equal the FP64 ones, ties broken by id.
//
// Tiling: a workgroup of 8 waves owns 128 query rows (16 per wave, their
// operands in registers as float4 per 16 k) and walks all base points in
// tiles of 64, double-buffered through LDS transposed ([point][k]).  Every
// tile gives a wave 16 x 64 distances; a distance enters its row's list only
// if it beats the row's current k-th best (kept in registers), so after the
// first tiles almost nothing is inserted.  Lists live in LDS, one per row,
// sorted by (dist, id); an insertion is done by the whole wave (lane i =
// list position i): the position is a ballot count and the tail shifts one
// lane up.  Points arrive in increasing id order, so a tie never displaces
// an earlier id.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "knn.h"

namespace rte {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kKnnRows = 128;  // rows per workgroup
constexpr int kKnnCols = 64;   // base points per tile

// |x|^2 per row, FP32, dims in order
__global__ void row_norms_kernel(const float* __restrict__ X, int64_t np, int d, int64_t stride,
                                 float* __restrict__ norms) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float* x = X + i * stride;
  float s = 0.f;
  for (int j = 0; j < d; ++j) s = __builtin_fmaf(x[j], x[j], s);
  norms[i] = s;
}

template <int S>  // S = d / 16 (float4 operands per lane)
__global__ __launch_bounds__(512) void knn_kernel(const float* __restrict__ X, int64_t np, int d,
                                                  int64_t stride, const float* __restrict__ norms,
                                                  int64_t row0, int64_t nq, int k,
                                                  int32_t* __restrict__ ids_out,
                                                  float* __restrict__ dists_out) {
  constexpr int LDB = 16 * S + 4;  // floats per point in LDS (padding breaks bank conflicts)
  __shared__ __attribute__((aligned(16))) float Bt[2][kKnnCols * LDB];
  __shared__ float Ld[kKnnRows * kKnnK];
  __shared__ int Li[kKnnRows * kKnnK];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int64_t mrow0 = row0 + (int64_t)blockIdx.x * kKnnRows + 16 * wave;  // wave's first row
  // lists empty
  for (int i = t; i < kKnnRows * kKnnK; i += 512) {
    Ld[i] = INFINITY;
    Li[i] = 0x7fffffff;
  }
  // A operands: row mrow0 + r, dims 16 s + 4 g .. + 3
  float4 A[S];
  {
    const int64_t row = mrow0 + r;
    const bool ok = row < row0 + nq;
#pragma unroll
    for (int s = 0; s < S; ++s)
      A[s] = ok ? *reinterpret_cast<const float4*>(X + row * stride + 16 * s + 4 * g)
                : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // lane (r, g) holds distances of rows 4 g + v (v < 4): their norms and thresholds
  float qn[4], thr[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t row = mrow0 + 4 * g + v;
    qn[v] = row < row0 + nq ? norms[row] : 0.f;
    thr[v] = INFINITY;
  }
  // tile loader: 64 points x d floats = 16 S float4 per point; thread t moves
  // float4 chunks t + 512 i
  constexpr int CH = kKnnCols * 4 * S;  // float4 chunks per tile
  constexpr int PER = (CH + 511) / 512;
  float4 pre[PER];
  auto chunk_src = [&](int64_t n0, int c, bool& ok) -> const float4* {
    const int p = c / (4 * S), q4 = c % (4 * S);
    ok = n0 + p < np;
    return reinterpret_cast<const float4*>(X + (n0 + p) * stride + 4 * q4);
  };
  const int64_t ntile = (np + kKnnCols - 1) / kKnnCols;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = t + 512 * i;
    bool ok = false;
    const float4* src = c < CH ? chunk_src(0, c, ok) : nullptr;
    pre[i] = (c < CH && ok) ? *src : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = t + 512 * i;
    if (c < CH) {
      const int p = c / (4 * S), q4 = c % (4 * S);
      *reinterpret_cast<float4*>(&Bt[0][p * LDB + 4 * q4]) = pre[i];
    }
  }
  __syncthreads();
  for (int64_t tile = 0; tile < ntile; ++tile) {
    const int64_t n0 = tile * kKnnCols;
    const int buf = (int)(tile & 1);
    const int64_t nn = tile + 1 < ntile ? n0 + kKnnCols : 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int c = t + 512 * i;
      bool ok = false;
      const float4* src = c < CH ? chunk_src(nn, c, ok) : nullptr;
      pre[i] = (c < CH && ok) ? *src : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    floatx4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float a4[4] = {A[s].x, A[s].y, A[s].z, A[s].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 b4 = *reinterpret_cast<const float4*>(&Bt[buf][(16 * j + r) * LDB + 16 * s + 4 * g]);
        const float b[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[c], b[c], acc[j], 0, 0, 0);
      }
    }
    // distances: lane holds rows 4 g + v, points n0 + 16 j + r
    float xn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + 16 * j + r;
      xn[j] = col < np ? norms[col] : 0.f;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = mrow0 + 4 * g + v;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t col = n0 + 16 * j + r;
        const float dist = __builtin_fmaf(-2.f, acc[j][v], qn[v] + xn[j]);
        const bool pass = col < np && col != row && dist < thr[v];
        uint64_t m = __ballot(pass);
        while (m) {  // wave-uniform: one insertion per candidate, lowest lane first
          const int src = __ffsll((unsigned long long)m) - 1;
          m &= m - 1;
          const float cd = __shfl(dist, src, 64);
          const int ci = (int)(n0 + 16 * j + (src & 15));
          const int lrow = 16 * wave + 4 * (src >> 4) + v;  // row in the workgroup
          float* ld = Ld + lrow * kKnnK;
          int* li = Li + lrow * kKnnK;
          const float ed = ld[lane];
          const int ei = li[lane];
          const bool less = ed < cd || (ed == cd && ei < ci);
          const int pos = __popcll(__ballot(less));
          const float pd = __shfl_up(ed, 1, 64);
          const int pi = __shfl_up(ei, 1, 64);
          if (pos < kKnnK && lane >= pos) {
            ld[lane] = lane == pos ? cd : pd;
            li[lane] = lane == pos ? ci : pi;
          }
          // the row's k-th best, to the 16 lanes holding the row
          const float kth = __shfl(lane == pos ? cd : (lane > pos ? pd : ed), k - 1, 64);
          if ((lane >> 4) == (src >> 4)) thr[v] = kth;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int c = t + 512 * i;
      if (c < CH) {
        const int p = c / (4 * S), q4 = c % (4 * S);
        *reinterpret_cast<float4*>(&Bt[buf ^ 1][p * LDB + 4 * q4]) = pre[i];
      }
    }
    __syncthreads();
  }
  // lists out: row lrow = wave's 16 rows, lanes = positions
  for (int lr = 0; lr < 16; ++lr) {
    const int lrow = 16 * wave + lr;
    const int64_t row = mrow0 + lr;
    if (row < row0 + nq && lane < k) {
      ids_out[(row - row0) * k + lane] = Li[lrow * kKnnK + lane];
      if (dists_out) dists_out[(row - row0) * k + lane] = Ld[lrow * kKnnK + lane];
    }
  }
}

hipError_t launch_row_norms(const float* X, int64_t np, int d, int64_t stride, float* norms,
                            hipStream_t st) {
  hipLaunchKernelGGL(row_norms_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, X,
                     np, d, stride, norms);
  return hipGetLastError();
}

hipError_t launch_knn(const float* X, int64_t np, int d, int64_t stride, const float* norms,
                      int64_t row0, int64_t nq, int k, int32_t* ids, float* dists,
                      hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nq + kKnnRows - 1) / kKnnRows));
#define KNN(S)                                                                                    \
  hipLaunchKernelGGL(knn_kernel<S>, grid, dim3(512), 0, st, X, np, d, stride, norms, row0, nq, k, \
                     ids, dists)
  switch (d / 16) {
    case 2: KNN(2); break;
    case 4: KNN(4); break;
    case 6: KNN(6); break;
    case 8: KNN(8); break;
    default: return hipErrorInvalidValue;
  }
#undef KNN
  return hipGetLastError();
}

}  // namespace rte
