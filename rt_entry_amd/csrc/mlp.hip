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
#include "entry.h"
#include "entry_impl.h"
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

// ---- v_mfma_f32_16x16x32_bf16 on gfx950 (training GEMMs and inference) ----
// Operand layout of the 16x16x32 MFMA used here: lane l supplies row (A) or
// column (B) l % 16 and the 8 consecutive k values 8 * (l / 16) + 0..7 of a
// 32-wide k step (A and B use the same k assignment, so only the order of
// the k sum is implied); lane l holds D[4 * (l / 16) + v][l % 16], v = 0..3.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ floatx4 mfma_k32(uint4 a, uint4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

enum Epi { EPI_RELU = 0, EPI_LINEAR = 1, EPI_RELU_BWD = 2, EPI_STORE = 3 };

// 64 x 32 (rows x k) tile of a bf16 matrix whose element (row, k) sits at
// P[row * srow + k * sk], for LDS T[64][40] ([row][k], rows 80 bytes apart so
// every 16-byte MFMA operand read is aligned).  KC: k is the contiguous
// dimension (sk == 1); thread t takes the 8 k values 8 * (t & 3) + 0..7 of
// row t >> 2 as one 16-byte load.  Otherwise rows are contiguous (srow ==
// 1); thread t takes rows 8 * (t & 7) + 0..7 of k value t >> 3 as one
// 16-byte load and spreads them over 8 LDS rows.  Runs that cross the
// matrix edge, or strides that break 16-byte alignment (vec false), are
// loaded element by element with zeros outside.
template <bool KC>
__device__ __forceinline__ uint4 tile_load(const uint16_t* __restrict__ P, int64_t srow,
                                           int64_t sk, int R, int K, int r0, int k0, int t,
                                           bool vec) {
  const int r = KC ? r0 + (t >> 2) : r0 + 8 * (t & 7);
  const int k = KC ? k0 + 8 * (t & 3) : k0 + (t >> 3);
  const bool full = KC ? (r < R && k + 8 <= K) : (r + 8 <= R && k < K);
  if (vec && full) return *reinterpret_cast<const uint4*>(P + (int64_t)r * srow + (int64_t)k * sk);
  uint16_t e[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int ri = KC ? r : r + i, ki = KC ? k + i : k;
    e[i] = (ri < R && ki < K) ? P[(int64_t)ri * srow + (int64_t)ki * sk] : (uint16_t)0;
  }
  uint4 v;
  memcpy(&v, e, 16);
  return v;
}
template <bool KC>
__device__ __forceinline__ void tile_store(uint16_t (*T)[40], int t, uint4 v) {
  if constexpr (KC) {
    *reinterpret_cast<uint4*>(&T[t >> 2][8 * (t & 3)]) = v;
  } else {
    uint16_t e[8];
    memcpy(e, &v, 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) T[8 * (t & 7) + i][t >> 3] = e[i];
  }
}

// C[M x N] = op(A)[M x K] . op(B)[K x N] with bf16 operands addressed by
// strides (element (m,k) of A at A[m*sam + k*sak], (k,n) of B at
// B[k*sbk + n*sbn]); AK / BK: k is A's / B's contiguous dimension (else m /
// n is).  Workgroup tile 64x64, four waves of 32x32 (2x2 tiles of
// v_mfma_f32_16x16x32_bf16), K staged through LDS 32 at a time with 16-byte
// loads, the next K tile's loads in flight while the current one is
// multiplied.  Epilogues:
//  RELU      out = relu(acc + bias[n]) (f32) and its bf16 copy
//  LINEAR    out = acc + bias[n]
//  RELU_BWD  out = acc * (act[m,n] > 0) (f32) and its bf16 copy
//  STORE     out = acc
template <int EPI, bool AK, bool BK>
__global__ __launch_bounds__(256) void gemm_bf16(const uint16_t* __restrict__ A, int64_t sam,
                                                 int64_t sak, const uint16_t* __restrict__ B,
                                                 int64_t sbk, int64_t sbn, int M, int N, int K,
                                                 const float* __restrict__ bias,
                                                 const float* __restrict__ act,
                                                 float* __restrict__ out,
                                                 uint16_t* __restrict__ out16) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];  // [m][k]
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][40];  // [n][k]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // 16-byte loads need the assumed contiguous stride and 16-byte alignment
  const bool va = ((uintptr_t)A & 15) == 0 && (AK ? sak == 1 && (sam & 7) == 0
                                                  : sam == 1 && (sak & 7) == 0);
  const bool vb = ((uintptr_t)B & 15) == 0 && (BK ? sbk == 1 && (sbn & 7) == 0
                                                  : sbn == 1 && (sbk & 7) == 0);
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint4 ra = tile_load<AK>(A, sam, sak, M, K, m0, 0, t, va);
  uint4 rb = tile_load<BK>(B, sbn, sbk, N, K, n0, 0, t, vb);
  tile_store<AK>(As, t, ra);
  tile_store<BK>(Bs, t, rb);
  __syncthreads();
  const int r = lane & 15, kq = 8 * (lane >> 4);
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bool more = k0 + 32 < K;
    if (more) {
      ra = tile_load<AK>(A, sam, sak, M, K, m0, k0 + 32, t, va);
      rb = tile_load<BK>(B, sbn, sbk, N, K, n0, k0 + 32, t, vb);
    }
    uint4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const uint4*>(&As[wm + 16 * i + r][kq]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const uint4*>(&Bs[wn + 16 * j + r][kq]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma_k32(a[i], b[j], acc[i][j]);
    __syncthreads();
    if (more) {
      tile_store<AK>(As, t, ra);
      tile_store<BK>(Bs, t, rb);
      __syncthreads();
    }
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

__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

// Hidden layers of the predictor for 32 rows per workgroup (four waves; the
// waves split the output columns, wave w owning 16-column blocks w, w+4, ...):
//   x0 = bf16((x - mean) * inv_std)           (as input_kernel)  -> LDS
//   H1 = bf16(relu(x0 W1 + b1))                                  -> LDS
//   H2 = bf16(relu(H1 W2 + b2))                                  -> global
// The A operands come from LDS; the W^T fragments of four 32-wide k steps
// are loaded from global memory before their MFMAs, so one memory latency
// covers four steps.  CB = 16-column blocks per wave (columns >= d1 / d2 are
// computed but never stored).  Input rows are ldx floats
// apart.  Requires d0 % 32 == 0, ldx % 4 == 0, d1 % 32 == 0 and d2 % 32 == 0;
// dynamic LDS 32 * max((d0 + 8) + (d1 + 8), d2 + 8) * 2 bytes.  The H2 tile
// leaves through LDS as whole rows, so every global store is 16 bytes wide
// and no store waits on another.
template <int CB>
struct WFrag {
  uint4 b[4][CB];  // W^T fragments of four 32-wide k steps
};

// W^T fragments of k steps k0 .. k0 + 127 for this wave's column blocks
template <int CB>
__device__ __forceinline__ void load_wfrag(WFrag<CB>& f, const uint16_t* __restrict__ WT, int N,
                                           int K, int k0, int wave, int r, int g) {
  // every load is issued, clamped to a valid address, so the group is one
  // run of loads with no branches between them: fragments of k steps >= K
  // are never multiplied (hidden_layer skips those steps) and columns >= N
  // give outputs that are never stored
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int kk = k0 + 32 * u + 8 * g;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int col = 16 * (wave + 4 * c) + r;
      f.b[u][c] = *reinterpret_cast<const uint4*>(WT + (int64_t)min(col, N - 1) * K +
                                                  min(kk, K - 8));
    }
  }
}

// acc = As[32 x K] . W^T fragments; `first` holds the fragments of k steps
// 0..3 (loaded by the caller ahead of time), later groups are loaded while
// the previous group multiplies.
template <int CB>
__device__ __forceinline__ void hidden_layer(const uint16_t* __restrict__ As, int lda, int K,
                                             const uint16_t* __restrict__ WT, int N, int wave,
                                             int r, int g, WFrag<CB>& first,
                                             floatx4 (&acc)[2][CB]) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  WFrag<CB> nxt;
  for (int k0 = 0; k0 < K; k0 += 128) {
    if (k0 + 128 < K) load_wfrag<CB>(nxt, WT, N, K, k0 + 128, wave, r, g);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = k0 + 32 * u + 8 * g;
      if (k0 + 32 * u < K) {
        uint4 a[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
          a[rb] = *reinterpret_cast<const uint4*>(&As[(16 * rb + r) * lda + kk]);
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) acc[rb][c] = mfma_k32(a[rb], first.b[u][c], acc[rb][c]);
      }
    }
    if (k0 + 128 < K) first = nxt;
  }
}

struct HiddenArgs {
  const float* X;  // input rows, ldx floats apart
  int64_t ldx;
  int64_t n;
  int d0;
  const float* mean;
  const float* inv;
  const uint16_t* W1T;
  const float* b1;
  int d1;
  const uint16_t* W2T;
  const float* b2;
  int d2;
  uint16_t* H2;  // n x d2 bf16
};

// rows 32 blk .. 32 blk + 31 of the hidden layers (the workgroup's 256
// threads; lds: dynamic LDS of the size stated above)
template <int CB>
__device__ __forceinline__ void hidden_rows(const HiddenArgs& a, int64_t blk, uint16_t* lds) {
  const float* __restrict__ X = a.X;
  const int64_t ldx = a.ldx, n = a.n;
  const int d0 = a.d0, d1 = a.d1, d2 = a.d2;
  const float* __restrict__ mean = a.mean;
  const float* __restrict__ inv = a.inv;
  const uint16_t* __restrict__ W1T = a.W1T;
  const uint16_t* __restrict__ W2T = a.W2T;
  const float* __restrict__ b1 = a.b1;
  const float* __restrict__ b2 = a.b2;
  uint16_t* __restrict__ H2 = a.H2;
  const int ld0 = d0 + 8, ld1 = d1 + 8, ld2 = d2 + 8;
  uint16_t* Xs = lds;
  uint16_t* H1s = lds + 32 * ld0;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int64_t m0 = blk * 32;
  WFrag<CB> w;  // first W1 group, in flight while the input tile is built
  load_wfrag<CB>(w, W1T, d1, d0, 0, wave, r, g);
  // biases of this lane's columns, loaded once with the first group (columns
  // past the layer width read a valid entry and are never stored)
  float bias1[CB], bias2[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int col = 16 * (wave + 4 * c) + r;
    bias1[c] = b1[min(col, d1 - 1)];
    bias2[c] = b2[min(col, d2 - 1)];
  }
  {  // normalised bf16 input tile: 8 threads per row, float4 each; the
     // loads of four column steps (row, mean, 1/std) are all issued at
     // clamped, valid addresses before any is used, so four steps cost one
     // memory round trip instead of four
    const int row = t >> 3;
    const bool in = m0 + row < n;
    const float* xr = X + (in ? m0 + row : 0) * ldx;
    for (int cb = (t & 7) * 4; cb < d0; cb += 4 * 32) {
      float4 x[4], u[4], v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c4 = min(cb + 32 * i, d0 - 4);
        x[i] = *reinterpret_cast<const float4*>(xr + c4);
        u[i] = *reinterpret_cast<const float4*>(mean + c4);
        v[i] = *reinterpret_cast<const float4*>(inv + c4);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c4 = cb + 32 * i;
        if (c4 < d0) {
          uint2 o = make_uint2(0, 0);
          if (in) {
            o.x = pack2((x[i].x - u[i].x) * v[i].x, (x[i].y - u[i].y) * v[i].y);
            o.y = pack2((x[i].z - u[i].z) * v[i].z, (x[i].w - u[i].w) * v[i].w);
          }
          *reinterpret_cast<uint2*>(&Xs[row * ld0 + c4]) = o;
        }
      }
    }
  }
  __syncthreads();
  floatx4 acc[2][CB];
  hidden_layer<CB>(Xs, ld0, d0, W1T, d1, wave, r, g, w, acc);
  load_wfrag<CB>(w, W2T, d2, d1, 0, wave, r, g);  // first W2 group, in flight during the H1 store
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int col = 16 * (wave + 4 * c) + r;
    if (col >= d1) continue;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        H1s[(16 * rb + 4 * g + v) * ld1 + col] = f2bf(fmaxf(acc[rb][c][v] + bias1[c], 0.f));
  }
  __syncthreads();
  hidden_layer<CB>(H1s, ld1, d1, W2T, d2, wave, r, g, w, acc);
  // H2 tile through LDS (the input and H1 tiles are dead once every wave is
  // past layer 2), then whole rows to global memory in 16-byte chunks
  __syncthreads();
  uint16_t* Hs = lds;
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int col = 16 * (wave + 4 * c) + r;
    if (col >= d2) continue;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Hs[(16 * rb + 4 * g + v) * ld2 + col] = f2bf(fmaxf(acc[rb][c][v] + bias2[c], 0.f));
  }
  __syncthreads();
  const int cpr = d2 >> 3;  // 16-byte chunks per row
  for (int e = t; e < 32 * cpr; e += 256) {
    const int row = e / cpr, ch = e - row * cpr;
    if (m0 + row < n)
      *reinterpret_cast<uint4*>(H2 + (m0 + row) * d2 + 8 * ch) =
          *reinterpret_cast<const uint4*>(&Hs[row * ld2 + 8 * ch]);
  }
}

template <int CB>
__global__ __launch_bounds__(256) void mlp_hidden_kernel(HiddenArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
  hidden_rows<CB>(a, blockIdx.x, lds);
}

struct SelectArgs {
  const BoxNode* nodes;
  int n_nodes;
  const float* pq;  // PCA queries, stride floats apart
  int64_t nq;
  int stride;
  int32_t* hits;
};

// The hidden layers and the kd-tree entry selection of the same queries in
// one launch: workgroups below sel_blocks select the entries of 256
// queries each (the walk of select_entry_kernel), the others run the hidden
// layers of 32 rows each.  The selection (the fallback leaves) so runs
// beside the network with no second stream and no event between them.
// Dynamic LDS: the larger of the hidden layers' and kTopNodes * 32 bytes.
template <int CB>
__global__ __launch_bounds__(256) void mlp_hidden_select_kernel(HiddenArgs a, SelectArgs sa,
                                                                int sel_blocks) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
  if ((int)blockIdx.x < sel_blocks)
    select_entry_block(sa.nodes, sa.n_nodes, reinterpret_cast<BoxNode*>(lds), sa.pq, sa.nq,
                       sa.stride, sa.hits, (int64_t)blockIdx.x * 256);
  else
    hidden_rows<CB>(a, (int64_t)blockIdx.x - sel_blocks, lds);
}

// Softmax partial of one row over a set of columns: maximum logit, its lowest
// column, and sum exp(z - max).
struct SoftPart {
  float mx;
  float se;
  int ix;
};

__device__ __forceinline__ void soft_merge(float& mx, float& se, int& ix, float om, float os,
                                           int oi) {
  if (om > mx || (om == mx && oi < ix)) {
    const float t = mx;
    mx = om;
    ix = oi;
    om = t;
    const float u = se;
    se = os;
    os = u;
  }
  if (om > -INFINITY) se += os * __expf(om - mx);
}

// Online softmax update of one row with four logits of a tile (columns
// col0 + 16 j), without branches: running max, lowest argmax column and
// sum exp(z - max), the exponentials as 2^((z - max) log2(e)).  The
// difference is taken before the scaling so that the maximum's own term is
// exactly 1 (the confidence 1 / sum is then at most 1).
// The running max starts finite (-1e30, below every logit) with se = 0;
// logits are finite or -inf (padding columns), which contribute nothing.
__device__ __forceinline__ void soft_update(float& mx, float& se, int& ix, float z0, float z1,
                                            float z2, float z3, int col0) {
  constexpr float L2E = 1.4426950408889634f;
  const float tm = fmaxf(fmaxf(z0, z1), fmaxf(z2, z3));
  int tj = z2 == tm ? 32 : 48;
  tj = z1 == tm ? 16 : tj;
  tj = z0 == tm ? 0 : tj;
  asm volatile("" : "+v"(tj));  // keep the select below a v_cndmask, not a branch
  ix = tm > mx ? col0 + tj : ix;
  const float nm = fmaxf(mx, tm);
  se *= __builtin_amdgcn_exp2f((mx - nm) * L2E);
  mx = nm;
  se += (__builtin_amdgcn_exp2f((z0 - nm) * L2E) + __builtin_amdgcn_exp2f((z1 - nm) * L2E)) +
        (__builtin_amdgcn_exp2f((z2 - nm) * L2E) + __builtin_amdgcn_exp2f((z3 - nm) * L2E));
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void global_void_t;

// Direct-to-LDS copy of one 64-column tile of W3^T (64 x K bf16) and its 64
// biases.  The tile is stored in 16-byte units, unit(c, q) = c * CPR +
// (q ^ (c & SWM)) for column c and k chunk q (8 bf16): the XOR spreads the
// 16 columns that an MFMA operand read touches over distinct banks, without
// padding.  Each global_load_lds_dwordx4 of a wave fills 64 consecutive
// units; soff[i] is this lane's source offset for its i-th copy.  The copies
// have no register destination, so they are issued where they stand (a
// whole tile ahead of their use) and only the wait before the barrier at the
// end of the tile covers them.
template <int KS>
__device__ __forceinline__ void head_dma_tile(const uint16_t* WT, const float* bias, int col0,
                                              const int (&soff)[KS], uint4* dst, float* bdst,
                                              int wave, int lane) {
  constexpr int K = 32 * KS;
#pragma unroll
  for (int i = 0; i < KS; ++i)
    __builtin_amdgcn_global_load_lds((global_void_t*)(WT + (int64_t)col0 * K + soff[i]),
                                     (lds_void_t*)(dst + 64 * (4 * i + wave)), 16, 0, 0);
  if (wave == 0)
    __builtin_amdgcn_global_load_lds((global_void_t*)(bias + col0 + lane), (lds_void_t*)bdst,
                                     4, 0, 0);
}

// Logits of one 64-column tile for the wave's 32 rows: acc = bias + H W^T,
// the accumulators starting from the bias; the B fragments of k step s + PD
// are read while step s multiplies.
template <int KS, int PD>
__device__ __forceinline__ void head_logits(const uint4* Bt, const float* bt, int r, int g, int xr,
                                            const uint4 (&A)[2][KS], floatx4 (&acc)[2][4]) {
  constexpr int CPR = 4 * KS;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float bj = bt[16 * j + r];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) acc[rb][j] = floatx4{bj, bj, bj, bj};
  }
  uint4 b[PD + 1][4];
#pragma unroll
  for (int p = 0; p < PD && p < KS; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) b[p][j] = Bt[(16 * j + r) * CPR + ((4 * p + g) ^ xr)];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    if (s + PD < KS) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[(s + PD) % (PD + 1)][j] = Bt[(16 * j + r) * CPR + ((4 * (s + PD) + g) ^ xr)];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        acc[rb][j] = mfma_k32(A[rb][s], b[s % (PD + 1)][j], acc[rb][j]);
  }
}

// One tile's logits into the online softmax of the wave's 32 rows.  Columns
// >= N carry a -inf bias (zero weights), so their logits are -inf and drop
// out of max and sum.
__device__ __forceinline__ void head_softmax(const floatx4 (&acc)[2][4], int col0,
                                             float (&mx)[2][4], float (&se)[2][4],
                                             int (&ix)[2][4]) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      soft_update(mx[rb][v], se[rb][v], ix[rb][v], acc[rb][0][v], acc[rb][1][v], acc[rb][2][v],
                  acc[rb][3][v], col0);
}

template <int KS>
__device__ __forceinline__ void head_tile(const uint4* Bt, const float* bt, int r, int g, int xr,
                                          const uint4 (&A)[2][KS], int n0, float (&mx)[2][4],
                                          float (&se)[2][4], int (&ix)[2][4]) {
  floatx4 acc[2][4];
  head_logits<KS, 1>(Bt, bt, r, g, xr, A, acc);
  head_softmax(acc, n0 + r, mx, se, ix);
}

// bf16 H2 fragments of the wave's 32 rows (from row m0), zero past n
template <int KS>
__device__ __forceinline__ void head_rows(const uint16_t* __restrict__ H, int64_t n, int64_t m0,
                                          int r, int g, uint4 (&A)[2][KS]) {
  constexpr int K = 32 * KS;
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int64_t row = m0 + 16 * rb + r;
#pragma unroll
    for (int s = 0; s < KS; ++s)
      A[rb][s] = row < n ? *reinterpret_cast<const uint4*>(H + row * K + 32 * s + 8 * g)
                         : make_uint4(0, 0, 0, 0);
  }
}

__device__ __forceinline__ void head_state_init(float (&mx)[2][4], float (&se)[2][4],
                                                int (&ix)[2][4]) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      mx[rb][v] = -1e30f;  // below every logit, finite (see soft_update)
      se[rb][v] = 0.f;
      ix[rb][v] = 0x7fffffff;
    }
}

// merge the 16 lanes of each row (lanes differing in bits 0..3) and write
// one SoftPart per (row, column group)
__device__ __forceinline__ void head_store(float (&mx)[2][4], float (&se)[2][4], int (&ix)[2][4],
                                           int64_t n, int64_t m0, int r, int g, int G, int grp,
                                           SoftPart* __restrict__ parts) {
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        const float om = __shfl_xor(mx[rb][v], o, 64);
        const float os = __shfl_xor(se[rb][v], o, 64);
        const int oi = __shfl_xor(ix[rb][v], o, 64);
        soft_merge(mx[rb][v], se[rb][v], ix[rb][v], om, os, oi);
      }
      const int64_t row = m0 + 16 * rb + 4 * g + v;
      if (r == 0 && row < n) {
        SoftPart p;
        p.mx = mx[rb][v];
        p.se = se[rb][v];
        p.ix = ix[rb][v];
        parts[row * G + grp] = p;
      }
    }
}

// Output layer fused with the softmax reduction, for 128 rows x one group of
// `cg` columns per workgroup (grid.x = column group, so consecutive
// workgroups, which land on different XCDs, read disjoint slices of W3).
// Wave w keeps the bf16 H2 fragments of its 32 rows in registers for the
// whole group; 64-column tiles of W3^T ([Npad][K] bf16, rows >= N zero,
// Npad a multiple of 64) and their biases are copied straight into two LDS
// buffers, the next tile's copy in flight while the current tile is
// multiplied (one wait and one barrier per tile; two separate buffers, so
// the compiler can tell that the copy into one does not touch the reads of
// the other).  The accumulators start from the bias, so each logit z = acc
// feeds an online (max, lowest argmax column, sum exp) per row directly;
// logits are never stored.  The 16 lanes sharing a row merge their states
// at the end and write one SoftPart per (row, group).  K = 32 * KS; static
// LDS 2 * 64 * K * 2 + 512 bytes; bias has Npad entries, -inf beyond N.
template <int KS>
__global__ __launch_bounds__(256, 2) void mlp_head_kernel(const uint16_t* __restrict__ H,
                                                          int64_t n,
                                                          const uint16_t* __restrict__ WT,
                                                          const float* __restrict__ bias, int N,
                                                          int cg, SoftPart* __restrict__ parts) {
  constexpr int K = 32 * KS;
  constexpr int CPR = K / 8;  // 16-byte units per column
  constexpr int SWM = (CPR < 16 ? CPR : 16) - 1;
  __shared__ uint4 Bt0[64 * CPR];
  __shared__ uint4 Bt1[64 * CPR];
  __shared__ float bias0[64];
  __shared__ float bias1[64];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int xr = r & SWM;  // (16 j + r) & SWM for every column block j
  const int grp = blockIdx.x, G = gridDim.x;
  const int c0 = grp * cg;
  const int c1 = min(N, c0 + cg);
  const int64_t m0 = (int64_t)blockIdx.y * 128 + 32 * wave;
  int soff[KS];  // element offset inside a tile of this lane's i-th copy
#pragma unroll
  for (int i = 0; i < KS; ++i) {
    const int u = 64 * (4 * i + wave) + lane;
    const int c = u / CPR;
    soff[i] = c * K + 8 * ((u % CPR) ^ (c & SWM));
  }
  const int ntile = (c1 - c0 + 63) >> 6;
  head_dma_tile<KS>(WT, bias, c0, soff, Bt0, bias0, wave, lane);
  uint4 A[2][KS];
  head_rows<KS>(H, n, m0, r, g, A);
  float mx[2][4], se[2][4];
  int ix[2][4];
  head_state_init(mx, se, ix);
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): the first tile is in LDS
  __syncthreads();
  // two tiles per trip, so each half names its buffers statically; a tile
  // past the group's last copies the group's first again (never read)
  for (int tile = 0; tile < ntile; tile += 2) {
    {
      const int n0 = c0 + 64 * tile;
      head_dma_tile<KS>(WT, bias, tile + 1 < ntile ? n0 + 64 : c0, soff, Bt1, bias1, wave, lane);
      head_tile<KS>(Bt0, bias0, r, g, xr, A, n0, mx, se, ix);
      __builtin_amdgcn_s_waitcnt(0x0f70);
      __syncthreads();
    }
    if (tile + 1 < ntile) {
      const int n0 = c0 + 64 * (tile + 1);
      head_dma_tile<KS>(WT, bias, tile + 2 < ntile ? n0 + 64 : c0, soff, Bt0, bias0, wave, lane);
      head_tile<KS>(Bt1, bias1, r, g, xr, A, n0, mx, se, ix);
      __builtin_amdgcn_s_waitcnt(0x0f70);
      __syncthreads();
    }
  }
  head_store(mx, se, ix, n, m0, r, g, G, grp, parts);
}

// Copy of one 64-column tile of W3^T and its 64 biases into one LDS buffer
// of the eight-wave kernel: the tile is 64 * CPR 16-byte units, copy i of
// wave w moves units 64 * (8 i + w) + lane (KS / 2 copies per wave), the
// biases follow the tile (wave 0).
template <int KS>
__device__ __forceinline__ void head8_dma_tile(const uint16_t* WT, const float* bias, int col0,
                                               const int (&soff)[KS / 2], uint4* dst, int wave,
                                               int lane) {
  constexpr int K = 32 * KS;
  constexpr int CPR = 4 * KS;
#pragma unroll
  for (int i = 0; i < KS / 2; ++i)
    __builtin_amdgcn_global_load_lds((global_void_t*)(WT + (int64_t)col0 * K + soff[i]),
                                     (lds_void_t*)(dst + 64 * (8 * i + wave)), 16, 0, 0);
  if (wave == 0)
    __builtin_amdgcn_global_load_lds((global_void_t*)(bias + col0 + lane),
                                     (lds_void_t*)(dst + 64 * CPR), 4, 0, 0);
}

// The output layer for 256 rows x one group of `cg` columns per workgroup of
// eight waves (KS = 2, 4, 8).  On this part a SIMD does not run one wave's
// MFMAs beside another wave's softmax arithmetic by itself when both waves
// reach them at once, so the work is skewed: iteration t computes the
// logits of tile t into one accumulator set and the softmax update of tile
// t - 1 from the other; waves 0-3 issue the MFMAs first and waves 4-7 the
// softmax first (a scheduling barrier keeps each order), and the two waves
// that share a SIMD (w and w + 4) keep its matrix unit and its VALU busy
// together.  Each 64-column tile (and its biases) is copied straight into
// one of two LDS buffers, the next tile's copy in flight during the
// current iteration.  Per (row, group) the result equals mlp_head_kernel's:
// the same MFMA order and the same softmax update sequence per lane.
// Static LDS 2 * (64 * K * 2 + 256) bytes; one workgroup per CU.
template <int KS>
__global__ __launch_bounds__(512, 1) void mlp_head8_kernel(const uint16_t* __restrict__ H,
                                                           int64_t n,
                                                           const uint16_t* __restrict__ WT,
                                                           const float* __restrict__ bias, int N,
                                                           int cg, SoftPart* __restrict__ parts) {
  static_assert(KS == 2 || KS == 4 || KS == 8, "eight-wave head: KS in {2, 4, 8}");
  constexpr int K = 32 * KS;
  constexpr int CPR = K / 8;
  constexpr int SWM = (CPR < 16 ? CPR : 16) - 1;
  __shared__ uint4 B0[64 * CPR + 16];  // tile, then its 64 biases
  __shared__ uint4 B1[64 * CPR + 16];
  const float* b0 = reinterpret_cast<const float*>(B0 + 64 * CPR);
  const float* b1 = reinterpret_cast<const float*>(B1 + 64 * CPR);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int xr = r & SWM;
  const int grp = blockIdx.x, G = gridDim.x;
  const int c0 = grp * cg;
  const int c1 = min(N, c0 + cg);
  const int64_t m0 = (int64_t)blockIdx.y * 256 + 32 * wave;
  const bool mfma_first = wave < 4;
  int soff[KS / 2];
#pragma unroll
  for (int i = 0; i < KS / 2; ++i) {
    const int u = 64 * (8 * i + wave) + lane;
    const int c = u / CPR;
    soff[i] = c * K + 8 * ((u % CPR) ^ (c & SWM));
  }
  const int ntile = (c1 - c0 + 63) >> 6;
  head8_dma_tile<KS>(WT, bias, c0, soff, B0, wave, lane);
  uint4 A[2][KS];
  head_rows<KS>(H, n, m0, r, g, A);
  float mx[2][4], se[2][4];
  int ix[2][4];
  head_state_init(mx, se, ix);
  // acc1 starts as -inf logits: the softmax update of iteration 0 leaves the
  // state unchanged
  floatx4 acc0[2][4], acc1[2][4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc1[rb][j] = floatx4{-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): rows and the first tile
  __syncthreads();
  // two iterations per trip, so buffers and accumulator sets are named
  // statically; iteration tl reads buffer tl & 1 and copies tile tl + 1
  // into the other one (free: every wave passed the barrier after tl - 1)
#define RTE_HEAD8_ITER(BUF, BB, NBUF, ACC, PACC, TL)                                       \
  {                                                                                        \
    const int tl_ = (TL);                                                                  \
    if (tl_ + 1 < ntile) head8_dma_tile<KS>(WT, bias, c0 + 64 * (tl_ + 1), soff, NBUF, wave, lane); \
    if (mfma_first) {                                                                      \
      head_logits<KS, 2>(BUF, BB, r, g, xr, A, ACC);                                       \
      __builtin_amdgcn_sched_barrier(0);                                                   \
      head_softmax(PACC, c0 + 64 * (tl_ - 1) + r, mx, se, ix);                             \
    } else {                                                                               \
      head_softmax(PACC, c0 + 64 * (tl_ - 1) + r, mx, se, ix);                             \
      __builtin_amdgcn_sched_barrier(0);                                                   \
      head_logits<KS, 2>(BUF, BB, r, g, xr, A, ACC);                                       \
    }                                                                                      \
    __builtin_amdgcn_s_waitcnt(0x0f70);                                                    \
    __syncthreads();                                                                       \
  }
  for (int tl = 0; tl < ntile; tl += 2) {
    RTE_HEAD8_ITER(B0, b0, B1, acc0, acc1, tl)
    if (tl + 1 < ntile) RTE_HEAD8_ITER(B1, b1, B0, acc1, acc0, tl + 1)
  }
#undef RTE_HEAD8_ITER
  // the last tile's softmax update
  if ((ntile - 1) & 1)
    head_softmax(acc1, c0 + 64 * (ntile - 1) + r, mx, se, ix);
  else
    head_softmax(acc0, c0 + 64 * (ntile - 1) + r, mx, se, ix);
  head_store(mx, se, ix, n, m0, r, g, G, grp, parts);
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

// W [K][N] f32 -> WT [Npad][K] bf16, rows n >= N zero
__global__ void transpose_bf16_kernel(const float* __restrict__ W, int K, int N, int Npad,
                                      uint16_t* __restrict__ WT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)K * Npad) return;
  const int n = (int)(i / K), k = (int)(i % K);
  WT[i] = n < N ? f2bf(W[(int64_t)k * N + n]) : (uint16_t)0;
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
  DevBuf b3p;              // output bias padded to whole 64-column tiles
  DevBuf P2, parts;        // inference: H2 (bf16) and softmax partials
  bool wt_dirty = true;
  int pcap = 0;
  int head_slots = 0;      // head workgroups resident at once on the device
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

// gemm_bf16 with the operand layouts picked from the strides (k contiguous
// when its stride is 1)
template <int EPI>
void launch_gemm(dim3 grid, hipStream_t st, const uint16_t* A, int64_t sam, int64_t sak,
                 const uint16_t* B, int64_t sbk, int64_t sbn, int M, int N, int K,
                 const float* bias, const float* act, float* out, uint16_t* out16) {
  const bool ak = sak == 1, bk = sbk == 1;
#define RTE_GEMM(AK, BK)                                                                    \
  hipLaunchKernelGGL((rte::gemm_bf16<EPI, AK, BK>), grid, dim3(256), 0, st, A, sam, sak, B, \
                     sbk, sbn, M, N, K, bias, act, out, out16)
  if (ak && bk) RTE_GEMM(true, true);
  else if (ak) RTE_GEMM(true, false);
  else if (bk) RTE_GEMM(false, true);
  else RTE_GEMM(false, false);
#undef RTE_GEMM
}

// forward of n rows already converted to bf16 in X16; logits in Z
hipError_t forward(Mlp* m, int n, hipStream_t st) {
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  // H1 = relu(X W1 + b1): A = X16 (n x d0), B = W1 (d0 x d1, row-major)
  launch_gemm<rte::EPI_RELU>(gemm_grid(n, d1), st,
                     m->X16.as<uint16_t>(), (int64_t)d0, (int64_t)1, m->W16[0].as<uint16_t>(),
                     (int64_t)d1, (int64_t)1, n, d1, d0, m->b[0].as<float>(), nullptr,
                     m->H1.as<float>(), m->H1b.as<uint16_t>());
  launch_gemm<rte::EPI_RELU>(gemm_grid(n, d2), st,
                     m->H1b.as<uint16_t>(), (int64_t)d1, (int64_t)1, m->W16[1].as<uint16_t>(),
                     (int64_t)d2, (int64_t)1, n, d2, d1, m->b[1].as<float>(), nullptr,
                     m->H2.as<float>(), m->H2b.as<uint16_t>());
  launch_gemm<rte::EPI_LINEAR>(gemm_grid(n, d3), st,
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
  launch_gemm<rte::EPI_STORE>(gemm_grid(d2, d3), st,
                     m->H2b.as<uint16_t>(), (int64_t)1, (int64_t)d2, m->dZb.as<uint16_t>(),
                     (int64_t)d3, (int64_t)1, d2, d3, n, nullptr, nullptr, m->gW[2].as<float>(),
                     nullptr);
  hipLaunchKernelGGL(rte::colsum_kernel, dim3((d3 + 255) / 256), dim3(256), 0, st,
                     m->dZ.as<float>(), n, d3, m->gb[2].as<float>());
  // dH2 = (dZ W3^T) * (H2 > 0): B(k=j, n=i) = W3[i*d3 + j]
  launch_gemm<rte::EPI_RELU_BWD>(gemm_grid(n, d2), st,
                     m->dZb.as<uint16_t>(), (int64_t)d3, (int64_t)1, m->W16[2].as<uint16_t>(),
                     (int64_t)1, (int64_t)d3, n, d2, d3, nullptr, m->H2.as<float>(),
                     m->dH2.as<float>(), m->dH2b.as<uint16_t>());
  launch_gemm<rte::EPI_STORE>(gemm_grid(d1, d2), st,
                     m->H1b.as<uint16_t>(), (int64_t)1, (int64_t)d1, m->dH2b.as<uint16_t>(),
                     (int64_t)d2, (int64_t)1, d1, d2, n, nullptr, nullptr, m->gW[1].as<float>(),
                     nullptr);
  hipLaunchKernelGGL(rte::colsum_kernel, dim3((d2 + 255) / 256), dim3(256), 0, st,
                     m->dH2.as<float>(), n, d2, m->gb[1].as<float>());
  launch_gemm<rte::EPI_RELU_BWD>(gemm_grid(n, d1), st,
                     m->dH2b.as<uint16_t>(), (int64_t)d2, (int64_t)1, m->W16[1].as<uint16_t>(),
                     (int64_t)1, (int64_t)d2, n, d1, d2, nullptr, m->H1.as<float>(),
                     m->dH1.as<float>(), m->dH1b.as<uint16_t>());
  launch_gemm<rte::EPI_STORE>(gemm_grid(d0, d1), st,
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

}  // extern "C"

namespace {

// Prediction of n rows of X (row stride ldx floats).  The fallback leaves
// are `fallback`, or with sel_pq the kd-tree leaves of sel_pq (row stride
// sel_stride), selected by workgroups of the hidden layers' launch.
int predict_impl(rte_ctx* c, const float* X, int64_t ldx, int64_t n, const int32_t* fallback,
                 const float* sel_pq, int sel_stride, float min_conf, int32_t* leaf_out,
                 float* conf_out) {
  Mlp* m = c->mlp;
  MCHK(c, hipSetDevice(c->device), "hipSetDevice");
  const int d0 = m->dims[0], d1 = m->dims[1], d2 = m->dims[2], d3 = m->dims[3];
  // fused inference kernels: input and first hidden width multiples of 32
  // (first hidden <= 512), second hidden width 32 * 2^i <= 512, 16-byte
  // aligned input rows; other shapes take the training kernels
  const bool fast = d0 % 32 == 0 && d1 % 32 == 0 && d1 <= 512 &&
                    (d2 == 32 || d2 == 64 || d2 == 128 || d2 == 256 || d2 == 512) &&
                    ((uintptr_t)X & 15) == 0 && ldx % 4 == 0;
  if (sel_pq) MCHK(c, c->kd_hits.reserve((size_t)n * 4), "kd-tree leaves");
  if (!fast) {
    if (ldx != d0) return rte_fail(c, RTE_EUNSUP, "strided input needs the fused kernels");
    if (sel_pq) {
      MCHK(c, rte::launch_select_entry(c->nodes.as<rte::BoxNode>(), c->n_nodes, sel_pq, n,
                                       sel_stride, c->kd_hits.as<int32_t>(), c->stream),
           "select entry");
      fallback = c->kd_hits.as<int32_t>();
    }
    const int chunk = 16384;
    MCHK(c, reserve_batch(m, (int)std::min<int64_t>(n, chunk)), "batch buffers");
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
      // the output layer's rows are padded to whole 64-column tiles
      const int np = l == 2 ? (m->dims[3] + 63) / 64 * 64 : m->dims[l + 1];
      const int64_t nw = (int64_t)m->dims[l] * np;
      MCHK(c, m->WT16[l].reserve(nw * 2), "transposed weights");
      hipLaunchKernelGGL(rte::transpose_bf16_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256),
                         0, c->stream, m->W[l].as<float>(), m->dims[l], m->dims[l + 1], np,
                         m->WT16[l].as<uint16_t>());
    }
    MCHK(c, hipGetLastError(), "transpose weights");
    const int np = (d3 + 63) / 64 * 64;
    MCHK(c, m->b3p.reserve((size_t)np * 4), "padded bias");
    // -inf bias on the padding columns: their logits are -inf
    MCHK(c, hipMemsetD32Async((hipDeviceptr_t)m->b3p.p, 0xff800000u, (size_t)np, c->stream),
         "padded bias");
    MCHK(c, hipMemcpyAsync(m->b3p.p, m->b[2].p, (size_t)d3 * 4, hipMemcpyDeviceToDevice, c->stream),
         "padded bias");
    m->wt_dirty = false;
  }
  const int chunk = 65536;
  const int b0 = (int)std::min<int64_t>(n, chunk);
  if (b0 > m->pcap) {
    MCHK(c, m->P2.reserve((size_t)b0 * d2 * 2), "hidden activations");
    m->pcap = b0;
  }
  const int cbw = std::max((d1 + 63) / 64, (d2 + 63) / 64);
  const int cb = cbw <= 1 ? 1 : cbw <= 2 ? 2 : cbw <= 4 ? 4 : 8;
  const int ks = d2 / 32;
  const size_t lds_h = (size_t)32 * std::max((d0 + 8) + (d1 + 8), d2 + 8) * 2;
  const size_t lds_o = 0;  // the head kernel's LDS is static
  const void* hk = cb == 1   ? (const void*)rte::mlp_hidden_kernel<1>
                   : cb == 2 ? (const void*)rte::mlp_hidden_kernel<2>
                   : cb == 4 ? (const void*)rte::mlp_hidden_kernel<4>
                             : (const void*)rte::mlp_hidden_kernel<8>;
  // output layer: the eight-wave kernel (256 rows per workgroup) for K = 64,
  // 128 and 256; the four-wave kernel (128 rows) for K = 32 (one k step) and
  // K = 512 (whose row fragments fill the registers)
  const bool head8 = ks == 2 || ks == 4 || ks == 8;
  const int head_threads = head8 ? 512 : 256, rows_per_wg = head8 ? 256 : 128;
  const void* ok = ks == 1    ? (const void*)rte::mlp_head_kernel<1>
                   : ks == 2  ? (const void*)rte::mlp_head8_kernel<2>
                   : ks == 4  ? (const void*)rte::mlp_head8_kernel<4>
                   : ks == 8  ? (const void*)rte::mlp_head8_kernel<8>
                              : (const void*)rte::mlp_head_kernel<16>;
  MCHK(c, hipFuncSetAttribute(ok, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_o),
       "head kernel LDS size");
  MCHK(c, hipFuncSetAttribute(hk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_h),
       "hidden kernel LDS size");
  const size_t lds_hs = std::max(lds_h, (size_t)rte::kTopNodes * sizeof(rte::BoxNode));
  if (sel_pq) {
    const void* hsk = cb == 1   ? (const void*)rte::mlp_hidden_select_kernel<1>
                      : cb == 2 ? (const void*)rte::mlp_hidden_select_kernel<2>
                      : cb == 4 ? (const void*)rte::mlp_hidden_select_kernel<4>
                                : (const void*)rte::mlp_hidden_select_kernel<8>;
    MCHK(c, hipFuncSetAttribute(hsk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_hs),
         "hidden + selection kernel LDS size");
  }
  if (m->head_slots == 0) {
    int cus = 0, per_cu = 0;
    MCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device),
         "CU count");
    MCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ok, head_threads, lds_o),
         "head kernel occupancy");
    m->head_slots = std::max(1, cus) * std::max(1, per_cu);
  }
  // Output columns per head workgroup: split the columns into as many
  // groups as keep every workgroup of a chunk resident at once (a second
  // round of workgroups on some CUs would nearly double the kernel's time),
  // each group a whole number of 64-column tiles.
  const int ntile = (d3 + 63) / 64;
  const int64_t rblk = (b0 + rows_per_wg - 1) / rows_per_wg;
  const int gw = (int)std::max<int64_t>(1, std::min<int64_t>(ntile, m->head_slots / rblk));
  const int cg = (ntile + gw - 1) / gw * 64;
  const int G = (d3 + cg - 1) / cg;
  MCHK(c, m->parts.reserve((size_t)b0 * G * sizeof(rte::SoftPart)), "softmax partials");
  // with sel_pq the kd-tree leaves (the fallback) are selected in
  // workgroups of the hidden-layer launch
  if (sel_pq) fallback = c->kd_hits.as<int32_t>();
  for (int64_t s = 0; s < n; s += chunk) {
    const int64_t b = std::min<int64_t>(chunk, n - s);
    const float* Xs = X + s * ldx;
    const float* mean = m->in_mean.as<float>();
    const float* inv = m->in_inv.as<float>();
    const uint16_t* W1T = m->WT16[0].as<uint16_t>();
    const uint16_t* W2T = m->WT16[1].as<uint16_t>();
    const uint16_t* W3T = m->WT16[2].as<uint16_t>();
    uint16_t* H2 = m->P2.as<uint16_t>();
    rte::SoftPart* parts = m->parts.as<rte::SoftPart>();
    const rte::HiddenArgs ha{Xs, ldx, b, d0, mean, inv, W1T, m->b[0].as<float>(), d1,
                             W2T, m->b[1].as<float>(), d2, H2};
    const unsigned nh = (unsigned)((b + 31) / 32);
    if (sel_pq) {
      const rte::SelectArgs sa{c->nodes.as<rte::BoxNode>(), c->n_nodes, sel_pq + s * sel_stride,
                               b, sel_stride, c->kd_hits.as<int32_t>() + s};
      const int sel_blocks = (int)((b + 255) / 256);
      const dim3 gs(nh + (unsigned)sel_blocks);
#define HIDSEL(CB)                                                                          \
  hipLaunchKernelGGL(rte::mlp_hidden_select_kernel<CB>, gs, dim3(256), lds_hs, c->stream, ha, \
                     sa, sel_blocks)
      switch (cb) {
        case 1: HIDSEL(1); break;
        case 2: HIDSEL(2); break;
        case 4: HIDSEL(4); break;
        default: HIDSEL(8); break;
      }
#undef HIDSEL
    } else {
#define HID(CB) \
  hipLaunchKernelGGL(rte::mlp_hidden_kernel<CB>, dim3(nh), dim3(256), lds_h, c->stream, ha)
      switch (cb) {
        case 1: HID(1); break;
        case 2: HID(2); break;
        case 4: HID(4); break;
        default: HID(8); break;
      }
#undef HID
    }
    const dim3 go((unsigned)G, (unsigned)((b + rows_per_wg - 1) / rows_per_wg));
#define HEAD(KERNEL)                                                                            \
  hipLaunchKernelGGL(KERNEL, go, dim3(head_threads), lds_o, c->stream, H2, b, W3T,               \
                     m->b3p.as<float>(), d3, cg, parts)
    switch (ks) {
      case 1: HEAD(rte::mlp_head_kernel<1>); break;
      case 2: HEAD(rte::mlp_head8_kernel<2>); break;
      case 4: HEAD(rte::mlp_head8_kernel<4>); break;
      case 8: HEAD(rte::mlp_head8_kernel<8>); break;
      default: HEAD(rte::mlp_head_kernel<16>); break;
    }
#undef HEAD
    hipLaunchKernelGGL(rte::predict_combine, dim3((unsigned)((b + 255) / 256)), dim3(256), 0,
                       c->stream, parts, G, b, fallback ? fallback + s : nullptr, min_conf,
                       leaf_out + s, conf_out ? conf_out + s : nullptr);
    MCHK(c, hipGetLastError(), "predict");
  }
  return RTE_OK;
}

}  // namespace

extern "C" {

int rte_mlp_predict(rte_ctx* c, const float* X, int64_t n, const int32_t* fallback,
                    float min_conf, int32_t* leaf_out, float* conf_out) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if ((n > 0 && (!X || !leaf_out)) || n < 0)
    return rte_fail(c, RTE_EINVAL, "bad predict arguments");
  if (n == 0) return RTE_OK;
  return predict_impl(c, X, c->mlp->dims[0], n, fallback, nullptr, 0, min_conf, leaf_out,
                      conf_out);
}

int rte_predict_entry(rte_ctx* c, const float* PQ, int64_t nq, int stride, float min_conf,
                      int32_t* leaf_out, float* conf_out) {
  if (!c) return RTE_EINVAL;
  if (!c->mlp) return rte_fail(c, RTE_ESTATE, "rte_mlp_create not called");
  if (!c->n_nodes) return rte_fail(c, RTE_ESTATE, "rte_set_leaves not called");
  if ((nq > 0 && (!PQ || !leaf_out)) || nq < 0 || stride < std::max(3, c->mlp->dims[0]))
    return rte_fail(c, RTE_EINVAL, "bad predict_entry arguments (stride %d)", stride);
  if (nq == 0) return RTE_OK;
  return predict_impl(c, PQ, stride, nq, nullptr, PQ, stride, min_conf, leaf_out, conf_out);
}

}  // extern "C"
