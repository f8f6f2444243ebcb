// PCA query projection PQ = Q R[:, :d_out] - (mu R[:, :d_out]) as an FP32
// MFMA GEMM (v_mfma_f32_16x16x4_f32) on gfx950.
//
// Replaces the cuBLAS call of Graph::Search (graph.cu:262-263 ->
// matrixMultiply, matrix.cu:39-69, alpha = 1, beta = -1 against the row
// mu.R replicated by replicateVector, matrix.cu:142-157): the replicated
// operand is folded into the epilogue as a per-column bias, so no
// replication kernel and no extra nq x d_out buffer exist.
//
// Tiling: a workgroup computes a 64 x 64 output tile (four waves of 16 rows,
// each four 16x16 MFMA accumulators).  K is processed in chunks of 128: the
// chunk of R's 64 columns is staged into LDS transposed ([col][k], one
// float4 per operand read); its loads and each lane's Q fragments of the
// chunk (one float4 per 16 k) are all issued before any is used, so a chunk
// costs one memory round trip.
// Inside a 16-wide k step the MFMA's k index g maps to k0 + 4g + c for
// component c of the lane's float4; that only permutes the k sum.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "project.h"

namespace rte {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kProjChunk = 128;
constexpr int kProjLds = kProjChunk + 4;  // row pitch of the transposed chunk (floats)

__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ Q, int64_t nq,
                                                      int d, const float* __restrict__ R,
                                                      int d_out, const float* __restrict__ bias,
                                                      float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) float Rt[64 * kProjLds];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  const int col0 = blockIdx.y * 64;
  const int r16 = lane & 15;  // A row / B column inside a 16x16 tile
  const int g = lane >> 4;    // k group
  const int64_t arow = row0 + r16;
  const bool arow_ok = arow < nq;
  const bool vec_ok = (d & 3) == 0;
  floatx4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < d; kc += kProjChunk) {
    // R[kc .. kc+128)[col0 .. col0+64) -> Rt[col][k]: thread t handles column
    // t % 64 and k = kc + t / 64 + 4 j; all R and Q loads of the chunk are
    // issued before the first is used
    const int rc = t & 63, k_first = t >> 6;
    const int col = col0 + rc;
    float rv[kProjChunk / 4];
#pragma unroll
    for (int j = 0; j < kProjChunk / 4; ++j) {
      const int k = kc + k_first + 4 * j;
      rv[j] = (k < d && col < d_out) ? R[(int64_t)k * d_out + col] : 0.f;
    }
    float4 qa[kProjChunk / 16];
#pragma unroll
    for (int s = 0; s < kProjChunk / 16; ++s) {
      const int ka = kc + 16 * s + 4 * g;
      if (arow_ok && vec_ok && ka + 3 < d) {
        qa[s] = *reinterpret_cast<const float4*>(Q + arow * d + ka);
      } else {
        qa[s].x = (arow_ok && ka + 0 < d) ? Q[arow * d + ka + 0] : 0.f;
        qa[s].y = (arow_ok && ka + 1 < d) ? Q[arow * d + ka + 1] : 0.f;
        qa[s].z = (arow_ok && ka + 2 < d) ? Q[arow * d + ka + 2] : 0.f;
        qa[s].w = (arow_ok && ka + 3 < d) ? Q[arow * d + ka + 3] : 0.f;
      }
    }
#pragma unroll
    for (int j = 0; j < kProjChunk / 4; ++j) Rt[rc * kProjLds + k_first + 4 * j] = rv[j];
    __syncthreads();
    const int steps = min(kProjChunk, d - kc + 15) / 16;
#pragma unroll
    for (int s = 0; s < kProjChunk / 16; ++s) {
      if (s < steps) {
        const float a4[4] = {qa[s].x, qa[s].y, qa[s].z, qa[s].w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float4 b4 =
              *reinterpret_cast<const float4*>(&Rt[(16 * u + r16) * kProjLds + 16 * s + 4 * g]);
          const float b[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[c], b[c], acc[u], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // D layout: lane holds D[4*g + i][r16]
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int col = col0 + 16 * u + r16;
    if (col >= d_out) continue;
    const float bc = bias[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = row0 + 4 * g + i;
      if (row < nq) out[row * d_out + col] = acc[u][i] - bc;
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
