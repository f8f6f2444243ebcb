// PCA query projection PQ = Q R[:, :d_out] - (mu R[:, :d_out]) as an FP32
// MFMA GEMM (v_mfma_f32_16x16x4_f32) on gfx950.
//
// Replaces the cuBLAS call of Graph::Search (graph.cu:262-263 ->
// matrixMultiply, matrix.cu:39-69, alpha = 1, beta = -1 against the row
// mu.R replicated by replicateVector, matrix.cu:142-157): the replicated
// operand is folded into the epilogue as a per-column bias, so no
// replication kernel and no extra nq x d_out buffer exist.
//
// Tiling: a workgroup computes a 64 x 64 output tile with eight waves: wave
// w takes rows 16 (w & 3) .. 16 (w & 3) + 15 (four 16x16 MFMA accumulators)
// over one half of every 128-k chunk (k half w >> 2), and the two halves'
// sums are added at the end through LDS; each wave's dependent MFMA chain
// and its share of the loads are half of what one wave per row block would
// carry.  The chunk of R's 64 columns is staged into LDS transposed ([col][k],
// one float4 per operand read); its loads and each lane's Q fragments of the
// chunk (one float4 per 16 k) are all issued before any is used, so a chunk
// costs one memory round trip.  The output tile leaves through LDS as
// 16-byte row segments.
// Inside a 16-wide k step the MFMA's k index g maps to k0 + 4g + c for
// component c of the lane's float4; that and the split into two halves only
// reorder the k sum (the FP32 error bound of the GEMM is unchanged).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "project.h"

namespace rte {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kProjChunk = 128;
constexpr int kProjLds = kProjChunk + 4;  // row pitch of the transposed chunk (floats)

template <bool VEC>  // VEC: d % 4 == 0, query rows are read as float4
__global__ __launch_bounds__(512) void project_kernel(const float* __restrict__ Q, int64_t nq,
                                                      int d, const float* __restrict__ R,
                                                      int d_out, const float* __restrict__ bias,
                                                      float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) float Rt[64 * kProjLds];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int rw = wave & 3;    // row block of the wave
  const int half = wave >> 2; // k half of every chunk
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + rw) * 16;
  const int col0 = blockIdx.y * 64;
  const int r16 = lane & 15;  // A row / B column inside a 16x16 tile
  const int g = lane >> 4;    // k group
  const int64_t arow = row0 + r16;
  const bool arow_ok = arow < nq;
  const float* qrow = Q + (arow_ok ? arow : nq - 1) * d;  // clamped: always a valid row
  constexpr int kHalfSteps = kProjChunk / 32;  // 16-wide k steps per half chunk
  floatx4 acc[4];
  float bc[4];  // bias of the lane's four columns, loaded once up front
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
    bc[u] = bias[min(col0 + 16 * u + r16, d_out - 1)];
  }
  for (int kc = 0; kc < d; kc += kProjChunk) {
    // R[kc .. kc+128)[col0 .. col0+64) -> Rt[col][k]: thread t handles column
    // t % 64 and k = kc + t / 64 + 8 j.  Every load of the chunk (R and this
    // lane's Q fragments of its k half) is issued at a clamped, valid address
    // and masked afterwards, with no branches, so all of them are in flight
    // together and are consumed before the barrier.
    const int rc = t & 63, k_first = t >> 6;
    const int col = col0 + rc;
    const int colc = min(col, d_out - 1);
    float rv[kProjChunk / 8];
#pragma unroll
    for (int j = 0; j < kProjChunk / 8; ++j) {
      const int k = kc + k_first + 8 * j;
      rv[j] = R[(int64_t)min(k, d - 1) * d_out + colc];
    }
    const int kh = kc + half * (kProjChunk / 2);  // first k of this wave's half
    float4 qa[kHalfSteps];
#pragma unroll
    for (int s = 0; s < kHalfSteps; ++s) {
      const int ka = kh + 16 * s + 4 * g;
      if (VEC) {
        qa[s] = *reinterpret_cast<const float4*>(qrow + min(ka, d - 4));
      } else {
        qa[s].x = qrow[min(ka + 0, d - 1)];
        qa[s].y = qrow[min(ka + 1, d - 1)];
        qa[s].z = qrow[min(ka + 2, d - 1)];
        qa[s].w = qrow[min(ka + 3, d - 1)];
      }
    }
#pragma unroll
    for (int j = 0; j < kProjChunk / 8; ++j) {
      const int k = kc + k_first + 8 * j;
      Rt[rc * kProjLds + k_first + 8 * j] = (k < d && col < d_out) ? rv[j] : 0.f;
    }
#pragma unroll
    for (int s = 0; s < kHalfSteps; ++s) {
      const int ka = kh + 16 * s + 4 * g;
      qa[s].x = (arow_ok && ka + 0 < d) ? qa[s].x : 0.f;
      qa[s].y = (arow_ok && ka + 1 < d) ? qa[s].y : 0.f;
      qa[s].z = (arow_ok && ka + 2 < d) ? qa[s].z : 0.f;
      qa[s].w = (arow_ok && ka + 3 < d) ? qa[s].w : 0.f;
    }
    __syncthreads();
    const int steps = min(kHalfSteps, max(0, d - kh + 15) / 16);
#pragma unroll
    for (int s = 0; s < kHalfSteps; ++s) {
      if (s < steps) {
        const float a4[4] = {qa[s].x, qa[s].y, qa[s].z, qa[s].w};
        const int kl = half * (kProjChunk / 2) + 16 * s + 4 * g;  // k inside the chunk
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float4 b4 = *reinterpret_cast<const float4*>(&Rt[(16 * u + r16) * kProjLds + kl]);
          const float b[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[c], b[c], acc[u], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // the second k half's sums through LDS (the R chunk buffer is free after
  // the last barrier): wave 4 + w writes, wave w adds
  floatx4* part = reinterpret_cast<floatx4*>(Rt);  // [4 row blocks][4 u][64 lanes]
  if (half == 1) {
#pragma unroll
    for (int u = 0; u < 4; ++u) part[(rw * 4 + u) * 64 + lane] = acc[u];
  }
  __syncthreads();
  if (half == 0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] += part[(rw * 4 + u) * 64 + lane];
  }
  __syncthreads();
  // D layout: lane holds D[4*g + i][r16].  The 64 x 64 tile goes through LDS
  // and leaves as whole row segments of 16-byte stores: single-float stores,
  // each under its own bounds branch, were each made to wait for the
  // previous store.
  float* T = Rt;  // [64][kProjLds]
  if (half == 0) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        T[(16 * rw + 4 * g + i) * kProjLds + 16 * u + r16] = acc[u][i] - bc[u];
  }
  __syncthreads();
  const int64_t trow0 = (int64_t)blockIdx.x * 64;
  const bool vec_out = (d_out & 3) == 0;
  for (int e = t; e < 64 * 16; e += 512) {
    const int rr = e >> 4, c4 = 4 * (e & 15);
    const int64_t row = trow0 + rr;
    const int col = col0 + c4;
    if (row >= nq || col >= d_out) continue;
    const float4 v = *reinterpret_cast<const float4*>(&T[rr * kProjLds + c4]);
    float* o = out + row * d_out + col;
    if (vec_out && col + 3 < d_out) {
      *reinterpret_cast<float4*>(o) = v;
    } else {
      const float w4[4] = {v.x, v.y, v.z, v.w};
      for (int k = 0; k < 4 && col + k < d_out; ++k) o[k] = w4[k];
    }
  }
}

hipError_t launch_project(const float* Q, int64_t nq, int d, const float* R, int d_out,
                          const float* bias, float* out, hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  const int64_t row_blocks = (nq + 63) / 64;
  dim3 grid((unsigned)row_blocks, (unsigned)((d_out + 63) / 64));
  if (d % 4 == 0 && (reinterpret_cast<uintptr_t>(Q) & 15) == 0)
    hipLaunchKernelGGL(project_kernel<true>, grid, dim3(512), 0, st, Q, nq, d, R, d_out, bias, out);
  else
    hipLaunchKernelGGL(project_kernel<false>, grid, dim3(512), 0, st, Q, nq, d, R, d_out, bias,
                       out);
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
