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

// ---- search graph from kNN lists -----------------------------------------
// Candidate edges of u: forward u -> knn[u][j] (rank j, type 0) and reverse
// u -> v for every v with knn[v][j] == u, j < rev_rank (rank j, type 1); a
// pair (u, v) keeps its smallest (rank, type); row u keeps the `degree`
// candidates of smallest (rank, type, v), padded with np.

__global__ void rev_count_kernel(const int32_t* __restrict__ knn, int64_t np, int kf, int kr,
                                 int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * kr) return;
  const int64_t v = i / kr;
  const int j = (int)(i - v * kr);
  const int32_t u = knn[v * kf + j];
  if (u >= 0 && u < np && u != v) atomicAdd(&cnt[u], 1);
}

// exclusive scan in blocks of 1024 (block totals to `sums`), then the block
// totals by one workgroup, then the carry-in
__global__ __launch_bounds__(1024) void scan_blocks_kernel(const int32_t* __restrict__ in, int64_t n,
                                                          int64_t* __restrict__ out,
                                                          int64_t* __restrict__ sums) {
  __shared__ int64_t sh[1024];
  const int t = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 1024 + t;
  const int64_t x = i < n ? in[i] : 0;
  sh[t] = x;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int64_t y = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += y;
    __syncthreads();
  }
  if (i < n) out[i] = sh[t] - x;
  if (t == 1023) sums[blockIdx.x] = sh[t];
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(int64_t* __restrict__ sums, int64_t nb,
                                                        int64_t* __restrict__ total) {
  __shared__ int64_t sh[1024];
  const int t = threadIdx.x;
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
    const int64_t b = b0 + t;
    const int64_t x = b < nb ? sums[b] : 0;
    sh[t] = x;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int64_t y = t >= o ? sh[t - o] : 0;
      __syncthreads();
      sh[t] += y;
      __syncthreads();
    }
    if (b < nb) sums[b] = carry + sh[t] - x;
    const int64_t blk = sh[1023];
    __syncthreads();
    carry += blk;
  }
  if (t == 0) total[0] = carry;
}

__global__ void scan_add_kernel(int64_t* __restrict__ out, int64_t n,
                                const int64_t* __restrict__ sums, const int64_t* __restrict__ total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / 1024];
  if (i == n) out[n] = total[0];
}

__global__ void rev_fill_kernel(const int32_t* __restrict__ knn, int64_t np, int kf, int kr,
                                const int64_t* __restrict__ off, int32_t* __restrict__ cur,
                                int64_t* __restrict__ rev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * kr) return;
  const int64_t v = i / kr;
  const int j = (int)(i - v * kr);
  const int32_t u = knn[v * kf + j];
  if (u >= 0 && u < np && u != v) {
    const int slot = atomicAdd(&cur[u], 1);
    rev[off[u] + slot] = ((int64_t)j << 32) | (int64_t)v;  // key (rank, id)
  }
}

// (rank, type, id) order of two candidates
__device__ __forceinline__ bool cand_less(int r1, int t1, int v1, int r2, int t2, int v2) {
  return r1 < r2 || (r1 == r2 && (t1 < t2 || (t1 == t2 && v1 < v2)));
}

// one wave per node u
__global__ __launch_bounds__(256) void graph_rows_kernel(const int32_t* __restrict__ knn,
                                                         int64_t np, int kf, int degree,
                                                         const int64_t* __restrict__ off,
                                                         const int64_t* __restrict__ rev,
                                                         int32_t* __restrict__ graph) {
  const int lane = threadIdx.x & 63;
  const int64_t u = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (u >= np) return;  // whole waves, no workgroup barrier
  const int32_t fv = lane < kf ? knn[u * kf + lane] : -1;
  const bool fvalid = lane < kf && fv >= 0 && fv < np && fv != u;
  uint64_t fdrop = 0;  // forward positions displaced by a smaller-rank reverse edge
  constexpr int64_t kNone = 0x7fffffffffffffffLL;
  int64_t tk = kNone;  // lane i < degree: i-th best reverse key (rank << 32 | id)
  const int64_t r0 = off[u], r1 = off[u + 1];
  for (int64_t c0 = r0; c0 < r1; c0 += 64) {
    const int64_t idx = c0 + lane;
    int64_t key = idx < r1 ? rev[idx] : kNone;
    const int rj = (int)(key >> 32);
    const int rv = (int)(key & 0xffffffff);
    bool keep = key != kNone;
    for (int p = 0; p < kf; ++p) {  // is rv also a forward neighbour, at rank p?
      const int fp = __shfl(fv, p, 64);
      const bool match = keep && rv == fp;
      if (__ballot(match && p > rj)) fdrop |= 1ull << p;
      if (match && p <= rj) keep = false;
    }
    uint64_t m = __ballot(keep);
    while (m) {  // insert the survivors into the sorted top-`degree` list
      const int src = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int64_t ck = __shfl(key, src, 64);
      const bool less = lane < degree && tk < ck;
      const int pos = __popcll(__ballot(less));
      const int64_t prev = __shfl_up(tk, 1, 64);
      if (pos < degree && lane >= pos && lane < degree) tk = lane == pos ? ck : prev;
    }
  }
  // merge: forward survivors (rank = position, type 0) and the reverse list
  const bool fs = fvalid && !((fdrop >> lane) & 1ull);
  const uint64_t fmask = __ballot(fs);
  const int nf_before = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(fmask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)fmask, 0u));
  const bool ts = lane < degree && tk != kNone;
  const uint64_t tmask = __ballot(ts);
  const int tr = (int)(tk >> 32), tv = (int)(tk & 0xffffffff);
  // output position of each surviving element = its rank in its own list +
  // the number of elements of the other list that precede it
  int fpos = nf_before, tpos = lane;
  for (int q = 0; q < 64; ++q) {
    if ((tmask >> q) & 1ull) {  // reverse element q against this lane's forward element
      const int qr = __shfl(tr, q, 64), qv = __shfl(tv, q, 64);
      if (fs && cand_less(qr, 1, qv, lane, 0, fv)) ++fpos;
    }
    if ((fmask >> q) & 1ull) {  // forward element q against this lane's reverse element
      const int qv = __shfl(fv, q, 64);
      if (ts && cand_less(q, 0, qv, tr, 1, tv)) ++tpos;
    }
  }
  int32_t* row = graph + u * degree;
  const int total = __popcll(fmask) + __popcll(tmask);
  if (lane < degree && lane >= total) row[lane] = (int32_t)np;
  if (fs && fpos < degree) row[fpos] = fv;
  if (ts && tpos < degree) row[tpos] = tv;
}

hipError_t launch_graph_from_knn(const int32_t* knn, int64_t np, int kf, int degree, int rev_rank,
                                 int32_t* graph, void* work, hipStream_t st) {
  const int kr = rev_rank < kf ? rev_rank : kf;
  // work: cnt (np) | cur (np) | off (np + 1) | sums | total | rev
  char* w = static_cast<char*>(work);
  int32_t* cnt = reinterpret_cast<int32_t*>(w);
  int32_t* cur = cnt + np;
  int64_t* off = reinterpret_cast<int64_t*>(w + ((2 * np * 4 + 255) & ~255LL));
  const int64_t nb = (np + 1023) / 1024;
  int64_t* sums = off + np + 1;
  int64_t* total = sums + nb;
  int64_t* rev = total + 1;
  hipError_t e = hipMemsetAsync(cnt, 0, 2 * np * sizeof(int32_t), st);
  if (e != hipSuccess) return e;
  const int64_t ne = np * (int64_t)kr;
  if (kr > 0) {
    hipLaunchKernelGGL(rev_count_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, knn,
                       np, kf, kr, cnt);
  }
  hipLaunchKernelGGL(scan_blocks_kernel, dim3((unsigned)nb), dim3(1024), 0, st, cnt, np, off, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, st, sums, nb, total);
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((np + 1 + 255) / 256)), dim3(256), 0, st, off,
                     np, sums, total);
  if (kr > 0) {
    hipLaunchKernelGGL(rev_fill_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, knn,
                       np, kf, kr, off, cur, rev);
  }
  hipLaunchKernelGGL(graph_rows_kernel, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, st, knn, np,
                     kf, degree, off, rev, graph);
  return hipGetLastError();
}

size_t graph_from_knn_work_bytes(int64_t np, int kf, int rev_rank) {
  const int kr = rev_rank < kf ? rev_rank : kf;
  const int64_t nb = (np + 1023) / 1024;
  return (size_t)((2 * np * 4 + 255) & ~255LL) + (size_t)(np + 1 + nb + 1) * 8 +
         (size_t)np * kr * 8;
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
