// Learned entry-leaf predictor: MLP d_in -> h1 -> h2 -> n_leaf (ReLU), bf16
// weights and activations into bf16 MFMA GEMMs with FP32 accumulation,
// trained on the GPU (softmax cross-entropy, Adam on FP32 master weights).
// No counterpart in the reference; it replaces the kd-tree leaf as the search
// entry when its softmax confidence reaches the caller's threshold.
//
// Numerics (mirrored by the FP64 oracle in the tests): the input, every
// activation fed to a GEMM and every weight are rounded to bf16 (round to
// nearest even); products accumulate in FP32; biases, the softmax and the
// optimiser run in FP32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "mlp.h"

namespace rte {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __host__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum Epi { EPI_RELU = 0, EPI_LINEAR = 1, EPI_RELU_BWD = 2, EPI_STORE = 3 };

// C[M x N] = op(A)[M x K] . op(B)[K x N] with bf16 operands addressed by
// strides (element (m,k) of A at A[m*sam + k*sak], (k,n) of B at
// B[k*sbk + n*sbn]).  Workgroup tile 64x64, four waves of 32x32 (2x2 tiles of
// v_mfma_f32_16x16x16_bf16), K staged through LDS 32 at a time.  Epilogues:
//  RELU      out = relu(acc + bias[n]) (f32) and its bf16 copy
//  LINEAR    out = acc + bias[n]
//  RELU_BWD  out = acc * (act[m,n] > 0) (f32) and its bf16 copy
//  STORE     out = acc
template <int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(const uint16_t* __restrict__ A, int64_t sam,
                                                 int64_t sak, const uint16_t* __restrict__ B,
                                                 int64_t sbk, int64_t sbn, int M, int N, int K,
                                                 const float* __restrict__ bias,
                                                 const float* __restrict__ act,
                                                 float* __restrict__ out,
                                                 uint16_t* __restrict__ out16) {
  __shared__ uint16_t As[64][32 + 4];  // [m][k]
  __shared__ uint16_t Bs[64][32 + 4];  // [n][k]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = t + 256 * e;  // 0..2047 over a 64x32 tile
      const int r = idx >> 5, c = idx & 31;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < M && gk < K) ? A[(int64_t)gm * sam + (int64_t)gk * sak] : (uint16_t)0;
      const int gn = n0 + r;
      Bs[r][c] = (gn < N && gk < K) ? B[(int64_t)gk * sbk + (int64_t)gn * sbn] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; kk += 16) {
      short4_t a[2], b[2];
      const int kq = kk + 4 * (lane >> 4);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + 16 * i + (lane & 15);
        a[i] = short4_t{(short)As[r][kq], (short)As[r][kq + 1], (short)As[r][kq + 2],
                        (short)As[r][kq + 3]};
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int c = wn + 16 * j + (lane & 15);
        b[j] = short4_t{(short)Bs[c][kq], (short)Bs[c][kq + 1], (short)Bs[c][kq + 2],
                        (short)Bs[c][kq + 3]};
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: lane holds D[4*(lane/16) + v][lane%16]
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        if (m >= M) continue;
        const int64_t o = (int64_t)m * N + n;
        float x = acc[i][j][v];
        if constexpr (EPI == EPI_RELU) {
          x = fmaxf(x + bias[n], 0.f);
          out[o] = x;
          out16[o] = f2bf(x);
        } else if constexpr (EPI == EPI_LINEAR) {
          out[o] = x + bias[n];
        } else if constexpr (EPI == EPI_RELU_BWD) {
          x = act[o] > 0.f ? x : 0.f;
          out[o] = x;
          out16[o] = f2bf(x);
        } else {
          out[o] = x;
        }
      }
    }
}

// Inference GEMM C = A . B^T with A [M][K] and B [N][K] bf16, both
// k-contiguous (K a multiple of 8).  Workgroup tile 128x128 (four waves of
// 64x64, 4x4 v_mfma_f32_16x16x16_bf16 tiles), K staged 32 at a time through
// LDS with 16-byte global loads and 8-byte operand reads.  Epilogues:
//  RELU16   out16 = bf16(relu(acc + bias[n]))
//  SOFTMAX  per row and 64-column wave tile: max of acc + bias, its lowest
//           column and sum exp(z - max), written as partials for
//           predict_combine (logits are never stored)
enum FastEpi { FEPI_RELU16 = 0, FEPI_SOFTMAX = 1 };

struct SoftPart {
  float mx;
  float se;
  int ix;
};

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_bf16(const uint16_t* __restrict__ A, int lda,
                                                    const uint16_t* __restrict__ B, int ldb,
                                                    int M, int N, int K,
                                                    const float* __restrict__ bias,
                                                    uint16_t* __restrict__ out16,
                                                    SoftPart* __restrict__ parts, int n_parts) {
  constexpr int LDSK = 40;  // 32 + 8 padding (row = 80 bytes)
  __shared__ __attribute__((aligned(16))) uint16_t As[128 * LDSK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[128 * LDSK];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * 128, n0 = blockIdx.x * 128;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = t + 256 * e;  // 512 chunks of 8 elements per 128x32 tile
      const int r = c >> 2, c8 = (c & 3) * 8;
      const int gk = k0 + c8;
      uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
      if (m0 + r < M && gk < K)
        va = *reinterpret_cast<const uint4*>(A + (int64_t)(m0 + r) * lda + gk);
      if (n0 + r < N && gk < K)
        vb = *reinterpret_cast<const uint4*>(B + (int64_t)(n0 + r) * ldb + gk);
      *reinterpret_cast<uint4*>(&As[r * LDSK + c8]) = va;
      *reinterpret_cast<uint4*>(&Bs[r * LDSK + c8]) = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; kk += 16) {
      const int kq = kk + 4 * (lane >> 4);
      short4_t a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = *reinterpret_cast<const short4_t*>(&As[(wm + 16 * i + (lane & 15)) * LDSK + kq]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[j] = *reinterpret_cast<const short4_t*>(&Bs[(wn + 16 * j + (lane & 15)) * LDSK + kq]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: lane holds D[4*(lane/16) + v][lane%16] of each 16x16 tile
  if constexpr (EPI == FEPI_RELU16) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (n >= N) continue;
        const float bn = bias[n];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
          if (m < M) out16[(int64_t)m * N + n] = f2bf(fmaxf(acc[i][j][v] + bn, 0.f));
        }
      }
  } else {
    float bn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn + 16 * j + (lane & 15);
      bn[j] = n < N ? bias[n] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // this lane's 4 columns of row m, then the 16 lanes holding the row
        float mx = -INFINITY;
        int ix = 0x7fffffff;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + wn + 16 * j + (lane & 15);
          const float z = n < N ? acc[i][j][v] + bn[j] : -INFINITY;
          if (z > mx) {
            mx = z;
            ix = n;
          }
        }
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
          const float om = __shfl_xor(mx, o, 64);
          const int oi = __shfl_xor(ix, o, 64);
          if (om > mx || (om == mx && oi < ix)) {
            mx = om;
            ix = oi;
          }
        }
        float se = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + wn + 16 * j + (lane & 15);
          if (n < N) se += expf(acc[i][j][v] + bn[j] - mx);
        }
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) se += __shfl_xor(se, o, 64);
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        if ((lane & 15) == 0 && m < M) {
          SoftPart p;
          p.mx = mx;
          p.se = se;
          p.ix = ix;
          parts[(int64_t)m * n_parts + blockIdx.x * 2 + (wn >> 6)] = p;
        }
      }
  }
}

// row-wise combination of the softmax partials: leaf = lowest column of the
// maximum logit, confidence = max softmax probability, kd-tree fallback
__global__ void predict_combine(const SoftPart* __restrict__ parts, int n_parts, int64_t M,
                                const int32_t* __restrict__ fallback, float min_conf,
                                int32_t* __restrict__ leaf, float* __restrict__ conf) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const SoftPart* p = parts + m * n_parts;
  float g = -INFINITY;
  int gi = 0x7fffffff;
  for (int t = 0; t < n_parts; ++t)
    if (p[t].mx > g || (p[t].mx == g && p[t].ix < gi)) {
      g = p[t].mx;
      gi = p[t].ix;
    }
  float s = 0.f;
  for (int t = 0; t < n_parts; ++t)
    if (p[t].mx > -INFINITY) s += p[t].se * expf(p[t].mx - g);
  const float pr = 1.f / s;
  if (conf) conf[m] = pr;
  leaf[m] = (pr >= min_conf || !fallback) ? gi : fallback[m];
}

// W [K][N] f32 -> WT [N][K] bf16
__global__ void transpose_bf16_kernel(const float* __restrict__ W, int K, int N,
                                      uint16_t* __restrict__ WT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)K * N) return;
  const int n = (int)(i / K), k = (int)(i % K);
  WT[i] = f2bf(W[(int64_t)k * N + n]);
}

__global__ void to_bf16_kernel(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// network input: bf16((x - mean[j]) * inv_std[j]) in FP32 arithmetic
__global__ void input_kernel(const float* __restrict__ x, int64_t n, int d,
                             const float* __restrict__ mean, const float* __restrict__ inv,
                             uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * d) return;
  const int j = (int)(i % d);
  const float c = x[i] - mean[j];
  y[i] = f2bf(c * inv[j]);
}

// Row-wise softmax cross-entropy: dZ = (softmax(z) - onehot(y)) / n and the
// per-row loss log(sum exp(z - max)) + max - z[y].
__global__ __launch_bounds__(256) void softmax_xent_kernel(const float* __restrict__ Z, int C,
                                                           const int32_t* __restrict__ y, int n,
                                                           float* __restrict__ dZ,
                                                           uint16_t* __restrict__ dZ16,
                                                           float* __restrict__ row_loss) {
  __shared__ float red[256];
  const int row = blockIdx.x;
  const float* z = Z + (int64_t)row * C;
  float mx = -INFINITY;
  for (int c = threadIdx.x; c < C; c += 256) mx = fmaxf(mx, z[c]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  mx = red[0];
  __syncthreads();
  float se = 0.f;
  for (int c = threadIdx.x; c < C; c += 256) se += expf(z[c] - mx);
  red[threadIdx.x] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  se = red[0];
  const int lab = min(max(y[row], 0), C - 1);  // labels are clamped into [0, C)
  const float inv = 1.f / (float)n;
  for (int c = threadIdx.x; c < C; c += 256) {
    const float p = expf(z[c] - mx) / se;
    const float g = (p - (c == lab ? 1.f : 0.f)) * inv;
    dZ[(int64_t)row * C + c] = g;
    dZ16[(int64_t)row * C + c] = f2bf(g);
  }
  if (threadIdx.x == 0) row_loss[row] = logf(se) + mx - z[lab];
}

// column sums of dZ (n x C): bias gradients
__global__ void colsum_kernel(const float* __restrict__ dZ, int n, int C, float* __restrict__ g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int r = 0; r < n; ++r) s += dZ[(int64_t)r * C + c];
  g[c] = s;
}

// Adam on FP32 masters; refreshes the bf16 copy
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ p16, int64_t n, float lr,
                            float b1, float b2, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  p[i] = pi;
  if (p16) p16[i] = f2bf(pi);
}

// argmax leaf and softmax maximum per row; falls back to the kd-tree leaf
// when the confidence is below min_conf
__global__ __launch_bounds__(256) void predict_kernel(const float* __restrict__ Z, int C,
                                                      const int32_t* __restrict__ fallback,
                                                      float min_conf, int32_t* __restrict__ leaf,
                                                      float* __restrict__ conf) {
  __shared__ float rmx[256];
  __shared__ int rix[256];
  const int row = blockIdx.x;
  const float* z = Z + (int64_t)row * C;
  float mx = -INFINITY;
  int ix = 0;
  for (int c = threadIdx.x; c < C; c += 256)
    if (z[c] > mx) {
      mx = z[c];
      ix = c;
    }
  rmx[threadIdx.x] = mx;
  rix[threadIdx.x] = ix;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float o = rmx[threadIdx.x + s];
      const int oi = rix[threadIdx.x + s];
      if (o > rmx[threadIdx.x] || (o == rmx[threadIdx.x] && oi < rix[threadIdx.x])) {
        rmx[threadIdx.x] = o;
        rix[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  mx = rmx[0];
  ix = rix[0];
  __syncthreads();
  float se = 0.f;
  for (int c = threadIdx.x; c < C; c += 256) se += expf(z[c] - mx);
  rmx[threadIdx.x] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) rmx[threadIdx.x] += rmx[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float p = 1.f / rmx[0];
    if (conf) conf[row] = p;
    leaf[row] = (p >= min_conf || !fallback) ? ix : fallback[row];
  }
}

struct Mlp {
  int dims[4];  // d_in, h1, h2, n_out
  DevBuf W[3], b[3], W16[3], mW[3], vW[3], mb[3], vb[3], gW[3], gb[3];
  DevBuf X16, H1, H1b, H2, H2b, Z, dZ, dZb, dH2, dH2b, dH1, dH1b, loss;
  DevBuf in_mean, in_inv;  // input normalisation (identity by default)
  DevBuf WT16[3];          // transposed bf16 weights for inference
  DevBuf P1, P2, parts;    // inference activations and softmax partials
  bool wt_dirty = true;
  int pcap = 0;
  int cap = 0;
  int step = 0;
};

void mlp_destroy(Mlp* m) { delete m; }

}  // namespace rte

rte_ctx::~rte_ctx() {
  rte::mlp_destroy(mlp);
  if (gexec) (void)hipGraphExecDestroy(gexec);
}

namespace {

using rte::Mlp;

int64_t wsize(const Mlp* m, int l) { return (int64_t)m->dims[l] * m->dims[l + 1]; }

hipError_t reserve_batch(Mlp* m, int n) {
  if (n <= m->cap) return hipSuccess;
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  hipError_t e;
#define R(buf, bytes) \
  if ((e = m->buf.reserve(bytes)) != hipSuccess) return e;
  R(X16, (size_t)n * d0 * 2);
  R(H1, (size_t)n * d1 * 4);
  R(H1b, (size_t)n * d1 * 2);
  R(H2, (size_t)n * d2 * 4);
  R(H2b, (size_t)n * d2 * 2);
  R(Z, (size_t)n * d3 * 4);
  R(dZ, (size_t)n * d3 * 4);
  R(dZb, (size_t)n * d3 * 2);
  R(dH2, (size_t)n * d2 * 4);
  R(dH2b, (size_t)n * d2 * 2);
  R(dH1, (size_t)n * d1 * 4);
  R(dH1b, (size_t)n * d1 * 2);
  R(loss, (size_t)n * 4);
#undef R
  m->cap = n;
  return hipSuccess;
}

dim3 gemm_grid(int M, int N) { return dim3((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64)); }

// forward of n rows already converted to bf16 in X16; logits in Z
hipError_t forward(Mlp* m, int n, hipStream_t st) {
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  // H1 = relu(X W1 + b1): A = X16 (n x d0), B = W1 (d0 x d1, row-major)
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_RELU>, gemm_grid(n, d1), dim3(256), 0, st,
                     m->X16.as<uint16_t>(), (int64_t)d0, (int64_t)1, m->W16[0].as<uint16_t>(),
                     (int64_t)d1, (int64_t)1, n, d1, d0, m->b[0].as<float>(), nullptr,
                     m->H1.as<float>(), m->H1b.as<uint16_t>());
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_RELU>, gemm_grid(n, d2), dim3(256), 0, st,
                     m->H1b.as<uint16_t>(), (int64_t)d1, (int64_t)1, m->W16[1].as<uint16_t>(),
                     (int64_t)d2, (int64_t)1, n, d2, d1, m->b[1].as<float>(), nullptr,
                     m->H2.as<float>(), m->H2b.as<uint16_t>());
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_LINEAR>, gemm_grid(n, d3), dim3(256), 0, st,
                     m->H2b.as<uint16_t>(), (int64_t)d2, (int64_t)1, m->W16[2].as<uint16_t>(),
                     (int64_t)d3, (int64_t)1, n, d3, d2, m->b[2].as<float>(), nullptr,
                     m->Z.as<float>(), nullptr);
  return hipGetLastError();
}

// backward after forward; gradients in gW/gb (mean cross-entropy)
hipError_t backward(Mlp* m, const int32_t* y, int n, hipStream_t st) {
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  hipLaunchKernelGGL(rte::softmax_xent_kernel, dim3(n), dim3(256), 0, st, m->Z.as<float>(), d3, y,
                     n, m->dZ.as<float>(), m->dZb.as<uint16_t>(), m->loss.as<float>());
  // gW3 = H2^T dZ: (d2 x n) . (n x d3); A(m=i,k=r) = H2b[r*d2+i]
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_STORE>, gemm_grid(d2, d3), dim3(256), 0, st,
                     m->H2b.as<uint16_t>(), (int64_t)1, (int64_t)d2, m->dZb.as<uint16_t>(),
                     (int64_t)d3, (int64_t)1, d2, d3, n, nullptr, nullptr, m->gW[2].as<float>(),
                     nullptr);
  hipLaunchKernelGGL(rte::colsum_kernel, dim3((d3 + 255) / 256), dim3(256), 0, st,
                     m->dZ.as<float>(), n, d3, m->gb[2].as<float>());
  // dH2 = (dZ W3^T) * (H2 > 0): B(k=j, n=i) = W3[i*d3 + j]
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_RELU_BWD>, gemm_grid(n, d2), dim3(256), 0, st,
                     m->dZb.as<uint16_t>(), (int64_t)d3, (int64_t)1, m->W16[2].as<uint16_t>(),
                     (int64_t)1, (int64_t)d3, n, d2, d3, nullptr, m->H2.as<float>(),
                     m->dH2.as<float>(), m->dH2b.as<uint16_t>());
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_STORE>, gemm_grid(d1, d2), dim3(256), 0, st,
                     m->H1b.as<uint16_t>(), (int64_t)1, (int64_t)d1, m->dH2b.as<uint16_t>(),
                     (int64_t)d2, (int64_t)1, d1, d2, n, nullptr, nullptr, m->gW[1].as<float>(),
                     nullptr);
  hipLaunchKernelGGL(rte::colsum_kernel, dim3((d2 + 255) / 256), dim3(256), 0, st,
                     m->dH2.as<float>(), n, d2, m->gb[1].as<float>());
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_RELU_BWD>, gemm_grid(n, d1), dim3(256), 0, st,
                     m->dH2b.as<uint16_t>(), (int64_t)d2, (int64_t)1, m->W16[1].as<uint16_t>(),
                     (int64_t)1, (int64_t)d2, n, d1, d2, nullptr, m->H1.as<float>(),
                     m->dH1.as<float>(), m->dH1b.as<uint16_t>());
  hipLaunchKernelGGL(rte::gemm_bf16<rte::EPI_STORE>, gemm_grid(d0, d1), dim3(256), 0, st,
                     m->X16.as<uint16_t>(), (int64_t)1, (int64_t)d0, m->dH1b.as<uint16_t>(),
                     (int64_t)d1, (int64_t)1, d0, d1, n, nullptr, nullptr, m->gW[0].as<float>(),
                     nullptr);
  hipLaunchKernelGGL(rte::colsum_kernel, dim3((d1 + 255) / 256), dim3(256), 0, st,
                     m->dH1.as<float>(), n, d1, m->gb[0].as<float>());
  return hipGetLastError();
}

hipError_t load_input(Mlp* m, const float* X, int64_t n, hipStream_t st) {
  const int64_t tot = n * m->dims[0];
  hipLaunchKernelGGL(rte::input_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, X,
                     n, m->dims[0], m->in_mean.as<float>(), m->in_inv.as<float>(),
                     m->X16.as<uint16_t>());
  return hipGetLastError();
}

// deterministic uniform init (splitmix64), He-uniform bound sqrt(6/fan_in)
void init_host(std::vector<float>& w, int fan_in, uint64_t& s) {
  const float bound = sqrtf(6.f / (float)fan_in);
  for (auto& x : w) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)((z >> 40) * (1.0 / 16777216.0));
    x = (2.f * u - 1.f) * bound;
  }
}

#define MCHK(c, expr, what)                                          \
  do {                                                               \
    hipError_t e_ = (expr);                                          \
    if (e_ != hipSuccess)                                            \
      return rte_fail(c, e_ == hipErrorOutOfMemory ? RTE_ENOMEM : RTE_EHIP, "%s: %s", what, \
                      hipGetErrorString(e_));                        \
  } while (0)

int upload_weights(rte_ctx* c, Mlp* m, const float* const* W, const float* const* b) {
  m->wt_dirty = true;
  for (int l = 0; l < 3; ++l) {
    const int64_t n = wsize(m, l);
    MCHK(c, hipMemcpyAsync(m->W[l].p, W[l], n * 4, hipMemcpyHostToDevice, c->stream), "copy W");
    MCHK(c, hipMemcpyAsync(m->b[l].p, b[l], m->dims[l + 1] * 4, hipMemcpyHostToDevice, c->stream),
         "copy b");
    hipLaunchKernelGGL(rte::to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, m->W[l].as<float>(), n, m->W16[l].as<uint16_t>());
    MCHK(c, hipGetLastError(), "to_bf16");
  }
  MCHK(c, hipStreamSynchronize(c->stream), "sync weights");
  return RTE_OK;
}

}  // namespace

extern "C" {

int rte_mlp_create(rte_ctx* c, int d_in, int h1, int h2, int n_out, uint64_t seed) {
  if (!c) return RTE_EINVAL;
  if (d_in < 1 || h1 < 1 || h2 < 1 || n_out < 2 || d_in > 8192 || h1 > 8192 || h2 > 8192 ||
      n_out > 1 << 20)
    return rte_fail(c, RTE_EINVAL, "bad mlp shape %d-%d-%d-%d", d_in, h1, h2, n_out);
  MCHK(c, hipSetDevice(c->device), "hipSetDevice");
  Mlp* m = new Mlp();
  m->dims[0] = d_in;
  m->dims[1] = h1;
  m->dims[2] = h2;
  m->dims[3] = n_out;
  for (int l = 0; l < 3; ++l) {
    const int64_t n = wsize(m, l);
    const int64_t nb = m->dims[l + 1];
    hipError_t e = hipSuccess;
    for (DevBuf* bf : {&m->W[l], &m->mW[l], &m->vW[l], &m->gW[l]})
      if (e == hipSuccess) e = bf->reserve(n * 4);
    if (e == hipSuccess) e = m->W16[l].reserve(n * 2);
    for (DevBuf* bf : {&m->b[l], &m->mb[l], &m->vb[l], &m->gb[l]})
      if (e == hipSuccess) e = bf->reserve(nb * 4);
    if (e == hipSuccess) e = hipMemsetAsync(m->mW[l].p, 0, n * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->vW[l].p, 0, n * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->mb[l].p, 0, nb * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->vb[l].p, 0, nb * 4, c->stream);
    if (e != hipSuccess) {
      delete m;
      return rte_fail(c, RTE_ENOMEM, "mlp allocation: %s", hipGetErrorString(e));
    }
  }
  {
    hipError_t e = m->in_mean.reserve((size_t)d_in * 4);
    if (e == hipSuccess) e = m->in_inv.reserve((size_t)d_in * 4);
    std::vector<float> zero(d_in, 0.f), one(d_in, 1.f);
    if (e == hipSuccess)
      e = hipMemcpy(m->in_mean.p, zero.data(), (size_t)d_in * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(m->in_inv.p, one.data(), (size_t)d_in * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      delete m;
      return rte_fail(c, RTE_ENOMEM, "mlp allocation: %s", hipGetErrorString(e));
    }
  }
  std::vector<float> hw[3], hb[3];
  uint64_t s = seed;
  for (int l = 0; l < 3; ++l) {
    hw[l].resize(wsize(m, l));
    init_host(hw[l], m->dims[l], s);
    hb[l].assign(m->dims[l + 1], 0.f);
  }
  const float* W[3] = {hw[0].data(), hw[1].data(), hw[2].data()};
  const float* B[3] = {hb[0].data(), hb[1].data(), hb[2].data()};
  rte::mlp_destroy(c->mlp);
  c->mlp = m;
  return upload_weights(c, m, W, B);
}

int rte_mlp_set_input_norm(rte_ctx* c, const float* mean, const float* inv_std) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!mean || !inv_std) return rte_fail(c, RTE_EINVAL, "null normalisation");
  Mlp* m = c->mlp;
  const size_t bytes = (size_t)m->dims[0] * 4;
  MCHK(c, hipMemcpyAsync(m->in_mean.p, mean, bytes, hipMemcpyHostToDevice, c->stream), "mean");
  MCHK(c, hipMemcpyAsync(m->in_inv.p, inv_std, bytes, hipMemcpyHostToDevice, c->stream), "inv");
  MCHK(c, hipStreamSynchronize(c->stream), "sync");
  return RTE_OK;
}

int rte_mlp_set(rte_ctx* c, const float* W1, const float* b1, const float* W2, const float* b2,
                const float* W3, const float* b3) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!W1 || !b1 || !W2 || !b2 || !W3 || !b3) return rte_fail(c, RTE_EINVAL, "null weights");
  const float* W[3] = {W1, W2, W3};
  const float* B[3] = {b1, b2, b3};
  c->mlp->step = 0;
  return upload_weights(c, c->mlp, W, B);
}

int rte_mlp_get(rte_ctx* c, float* W1, float* b1, float* W2, float* b2, float* W3, float* b3) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  Mlp* m = c->mlp;
  float* W[3] = {W1, W2, W3};
  float* B[3] = {b1, b2, b3};
  for (int l = 0; l < 3; ++l) {
    if (W[l])
      MCHK(c, hipMemcpyAsync(W[l], m->W[l].p, wsize(m, l) * 4, hipMemcpyDeviceToHost, c->stream),
           "read W");
    if (B[l])
      MCHK(c, hipMemcpyAsync(B[l], m->b[l].p, m->dims[l + 1] * 4, hipMemcpyDeviceToHost, c->stream),
           "read b");
  }
  MCHK(c, hipStreamSynchronize(c->stream), "sync");
  return RTE_OK;
}

int rte_mlp_grad(rte_ctx* c, const float* X, const int32_t* y, int n, float* gW1, float* gb1,
                 float* gW2, float* gb2, float* gW3, float* gb3, float* loss) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!X || !y || n < 1) return rte_fail(c, RTE_EINVAL, "bad batch");
  Mlp* m = c->mlp;
  MCHK(c, hipSetDevice(c->device), "hipSetDevice");
  MCHK(c, reserve_batch(m, n), "batch buffers");
  MCHK(c, load_input(m, X, n, c->stream), "input");
  MCHK(c, forward(m, n, c->stream), "forward");
  MCHK(c, backward(m, y, n, c->stream), "backward");
  float* G[3] = {gW1, gW2, gW3};
  float* B[3] = {gb1, gb2, gb3};
  for (int l = 0; l < 3; ++l) {
    if (G[l])
      MCHK(c, hipMemcpyAsync(G[l], m->gW[l].p, wsize(m, l) * 4, hipMemcpyDeviceToHost, c->stream),
           "read gW");
    if (B[l])
      MCHK(c, hipMemcpyAsync(B[l], m->gb[l].p, m->dims[l + 1] * 4, hipMemcpyDeviceToHost,
                             c->stream),
           "read gb");
  }
  if (loss) {
    std::vector<float> rl(n);
    MCHK(c, hipMemcpyAsync(rl.data(), m->loss.p, n * 4, hipMemcpyDeviceToHost, c->stream),
         "read loss");
    MCHK(c, hipStreamSynchronize(c->stream), "sync");
    double s = 0;
    for (float v : rl) s += v;
    *loss = (float)(s / n);
  }
  MCHK(c, hipStreamSynchronize(c->stream), "sync");
  return RTE_OK;
}

int rte_mlp_train_step(rte_ctx* c, const float* X, const int32_t* y, int n, float lr, float beta1,
                       float* loss) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!X || !y || n < 1) return rte_fail(c, RTE_EINVAL, "bad batch");
  if (!(beta1 >= 0.f && beta1 < 1.f)) return rte_fail(c, RTE_EINVAL, "beta1 outside [0,1)");
  Mlp* m = c->mlp;
  MCHK(c, hipSetDevice(c->device), "hipSetDevice");
  MCHK(c, reserve_batch(m, n), "batch buffers");
  MCHK(c, load_input(m, X, n, c->stream), "input");
  MCHK(c, forward(m, n, c->stream), "forward");
  MCHK(c, backward(m, y, n, c->stream), "backward");
  m->step += 1;
  m->wt_dirty = true;
  const float b2 = 0.999f;
  const float c1 = 1.f - powf(beta1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
  for (int l = 0; l < 3; ++l) {
    const int64_t nw = wsize(m, l), nb = m->dims[l + 1];
    hipLaunchKernelGGL(rte::adam_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                       c->stream, m->W[l].as<float>(), m->gW[l].as<float>(), m->mW[l].as<float>(),
                       m->vW[l].as<float>(), m->W16[l].as<uint16_t>(), nw, lr, beta1, b2, c1, c2);
    hipLaunchKernelGGL(rte::adam_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       c->stream, m->b[l].as<float>(), m->gb[l].as<float>(), m->mb[l].as<float>(),
                       m->vb[l].as<float>(), (uint16_t*)nullptr, nb, lr, beta1, b2, c1, c2);
  }
  MCHK(c, hipGetLastError(), "adam");
  if (loss) {
    std::vector<float> rl(n);
    MCHK(c, hipMemcpyAsync(rl.data(), m->loss.p, n * 4, hipMemcpyDeviceToHost, c->stream),
         "read loss");
    MCHK(c, hipStreamSynchronize(c->stream), "sync");
    double s = 0;
    for (float v : rl) s += v;
    *loss = (float)(s / n);
  }
  return RTE_OK;
}

int rte_mlp_predict(rte_ctx* c, const float* X, int64_t n, const int32_t* fallback,
                    float min_conf, int32_t* leaf_out, float* conf_out) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!X || !leaf_out || n < 0) return rte_fail(c, RTE_EINVAL, "bad predict arguments");
  if (n == 0) return RTE_OK;
  Mlp* m = c->mlp;
  MCHK(c, hipSetDevice(c->device), "hipSetDevice");
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  const int chunk = 16384;
  const int b0 = (int)std::min<int64_t>(n, chunk);
  MCHK(c, reserve_batch(m, b0), "batch buffers");
  const bool fast = d0 % 8 == 0 && d1 % 8 == 0 && d2 % 8 == 0;
  if (!fast) {  // generic kernels for shapes the 16-byte path cannot load
    for (int64_t s = 0; s < n; s += chunk) {
      const int b = (int)std::min<int64_t>(chunk, n - s);
      MCHK(c, load_input(m, X + s * d0, b, c->stream), "input");
      MCHK(c, forward(m, b, c->stream), "forward");
      hipLaunchKernelGGL(rte::predict_kernel, dim3(b), dim3(256), 0, c->stream,
                         m->Z.as<float>(), d3, fallback ? fallback + s : nullptr, min_conf,
                         leaf_out + s, conf_out ? conf_out + s : nullptr);
      MCHK(c, hipGetLastError(), "predict");
    }
    return RTE_OK;
  }
  if (m->wt_dirty) {
    for (int l = 0; l < 3; ++l) {
      const int64_t nw = wsize(m, l);
      MCHK(c, m->WT16[l].reserve(nw * 2), "transposed weights");
      hipLaunchKernelGGL(rte::transpose_bf16_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256),
                         0, c->stream, m->W[l].as<float>(), m->dims[l], m->dims[l + 1],
                         m->WT16[l].as<uint16_t>());
    }
    MCHK(c, hipGetLastError(), "transpose weights");
    m->wt_dirty = false;
  }
  const int n_parts = 2 * ((d3 + 127) / 128);
  if (b0 > m->pcap) {
    MCHK(c, m->P1.reserve((size_t)b0 * d1 * 2), "P1");
    MCHK(c, m->P2.reserve((size_t)b0 * d2 * 2), "P2");
    MCHK(c, m->parts.reserve((size_t)b0 * n_parts * sizeof(rte::SoftPart)), "parts");
    m->pcap = b0;
  } else {
    MCHK(c, m->parts.reserve((size_t)b0 * n_parts * sizeof(rte::SoftPart)), "parts");
  }
  for (int64_t s = 0; s < n; s += chunk) {
    const int b = (int)std::min<int64_t>(chunk, n - s);
    MCHK(c, load_input(m, X + s * d0, b, c->stream), "input");
    auto grid = [](int M_, int N_) {
      return dim3((unsigned)((N_ + 127) / 128), (unsigned)((M_ + 127) / 128));
    };
    hipLaunchKernelGGL(rte::gemm_nt_bf16<rte::FEPI_RELU16>, grid(b, d1), dim3(256), 0, c->stream,
                       m->X16.as<uint16_t>(), d0, m->WT16[0].as<uint16_t>(), d0, b, d1, d0,
                       m->b[0].as<float>(), m->P1.as<uint16_t>(), (rte::SoftPart*)nullptr, 0);
    hipLaunchKernelGGL(rte::gemm_nt_bf16<rte::FEPI_RELU16>, grid(b, d2), dim3(256), 0, c->stream,
                       m->P1.as<uint16_t>(), d1, m->WT16[1].as<uint16_t>(), d1, b, d2, d1,
                       m->b[1].as<float>(), m->P2.as<uint16_t>(), (rte::SoftPart*)nullptr, 0);
    hipLaunchKernelGGL(rte::gemm_nt_bf16<rte::FEPI_SOFTMAX>, grid(b, d3), dim3(256), 0, c->stream,
                       m->P2.as<uint16_t>(), d2, m->WT16[2].as<uint16_t>(), d2, b, d3, d2,
                       m->b[2].as<float>(), (uint16_t*)nullptr, m->parts.as<rte::SoftPart>(),
                       n_parts);
    hipLaunchKernelGGL(rte::predict_combine, dim3((unsigned)((b + 255) / 256)), dim3(256), 0,
                       c->stream, m->parts.as<rte::SoftPart>(), n_parts, (int64_t)b,
                       fallback ? fallback + s : nullptr, min_conf, leaf_out + s,
                       conf_out ? conf_out + s : nullptr);
    MCHK(c, hipGetLastError(), "predict");
  }
  return RTE_OK;
}

}  // extern "C"
