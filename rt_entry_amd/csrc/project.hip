// PCA query projection PQ = Q R[:, :d_out] - (mu R[:, :d_out]) as an FP32
// MFMA GEMM (v_mfma_f32_16x16x4_f32) on gfx950.
//
// Replaces the cuBLAS call of Graph::Search (graph.cu:262-263 ->
// matrixMultiply, matrix.cu:39-69, alpha = 1, beta = -1 against the row
// mu.R replicated by replicateVector, matrix.cu:142-157): the replicated
// operand is folded into the epilogue as a per-column bias, so no
// replication kernel and no extra nq x d_out buffer exist.
//
// Tiling: one wave computes a 16 x 64 output tile as four 16x16 MFMA
// accumulators; each lane loads Q as float4 along k (16 k per step), so one
// step issues 4 MFMAs per accumulator.  Inside a step the MFMA's k index g
// maps to k0 + 4g + c for component c; that only permutes the k sum.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "project.h"

namespace rte {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ Q, int64_t nq,
                                                      int d, const float* __restrict__ R,
                                                      int d_out, const float* __restrict__ bias,
                                                      float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  const int col0 = blockIdx.y * 64;
  if (row0 >= nq) return;
  const int r16 = lane & 15;  // A row / B column inside a 16x16 tile
  const int g = lane >> 4;    // k group
  const int64_t arow = row0 + r16;
  const bool arow_ok = arow < nq;
  floatx4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  const bool vec_ok = (d & 3) == 0;
  for (int k0 = 0; k0 < d; k0 += 16) {
    float a4[4];
    const int ka = k0 + 4 * g;
    if (arow_ok && vec_ok && ka + 3 < d) {
      const float4 v = *reinterpret_cast<const float4*>(Q + arow * d + ka);
      a4[0] = v.x;
      a4[1] = v.y;
      a4[2] = v.z;
      a4[3] = v.w;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) a4[c] = (arow_ok && ka + c < d) ? Q[arow * d + ka + c] : 0.f;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int kb = ka + c;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int col = col0 + 16 * t + r16;
        const float b = (kb < d && col < d_out) ? R[(int64_t)kb * d_out + col] : 0.f;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[c], b, acc[t], 0, 0, 0);
      }
    }
  }
  // D layout: lane holds D[4*g + i][r16]
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = col0 + 16 * t + r16;
    if (col >= d_out) continue;
    const float bc = bias[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = row0 + 4 * g + i;
      if (row < nq) out[row * d_out + col] = acc[t][i] - bc;
    }
  }
}

hipError_t launch_project(const float* Q, int64_t nq, int d, const float* R, int d_out,
                          const float* bias, float* out, hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  const int64_t row_blocks = (nq + 63) / 64;
  dim3 grid((unsigned)row_blocks, (unsigned)((d_out + 63) / 64));
  hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, st, Q, nq, d, R, d_out, bias, out);
  return hipGetLastError();
}

// bias[c] = sum_j mu[j] R[j, c]: the mu.R row of graph.cu:207-217, once per
// rte_set_pca, in FP32 with a fixed left-to-right order.
__global__ void mu_r_kernel(const float* __restrict__ mu, const float* __restrict__ R, int d,
                            int d_out, float* __restrict__ bias) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d_out) return;
  float s = 0.f;
  for (int j = 0; j < d; ++j) s = __builtin_fmaf(mu[j], R[(int64_t)j * d_out + c], s);
  bias[c] = s;
}

hipError_t launch_mu_r(const float* mu, const float* R, int d, int d_out, float* bias,
                       hipStream_t st) {
  hipLaunchKernelGGL(mu_r_kernel, dim3((d_out + 127) / 128), dim3(128), 0, st, mu, R, d, d_out,
                     bias);
  return hipGetLastError();
}

}  // namespace rte
