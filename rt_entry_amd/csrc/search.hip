// Greedy beam search on a proximity graph, one wave64 per query (gfx950).
//
// Restates GraphSearchKernel (graph_search.cuh:56-1964) for CDNA4:
//  - one wavefront owns one query; a 256-thread workgroup carries 4
//    independent queries and never uses a workgroup barrier;
//  - the beam L (N slots) and the expansion batch B (M = width * 2^shift
//    slots) live in VGPRs, slot s in lane s % 64 of register s / 64; the
//    reference's compare-exchange networks (bitonic sort of B
//    :888-919/:1424-1453, min-exchange :921-931/:1455-1467, half-cleaner
//    merge of L :933-958/:1469-1493) run as cross-lane exchanges, so every
//    pair sees exactly the comparison the reference makes;
//  - the visited-set filter (:4-43) lives in LDS, one region per wave;
//  - distances follow the reference association: each 32-lane half-wave
//    computes one point, lane l accumulates dims l+32j as an FMA chain, then
//    a shfl-down tree 16,8,4,2,1 (:686-878);
//  - neighbour ids that pass the filter are compacted before the gathers, so
//    loads are issued only for points that need a distance.
// Parity decisions P1-P7 match the CPU oracle (zeroed filter, all-ones first
// mask, snapshot filter test per row, no row read for ids >= np,
// barrier-separated substeps, REF/TUNED comparators).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "search.h"

namespace rte {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// graph_search.cuh:89-97
__constant__ uint32_t c_hash_consts[20] = {
    0x924ed183U, 0xd854fc0aU, 0xecf5e3b7U, 0x1bead407U, 0x28a30449U,
    0xbfc4d99fU, 0x715030e2U, 0xffcfb45bU, 0x6e4ce166U, 0xeb53c362U,
    0xa93c4f40U, 0xcecde0a4U, 0x0288592dU, 0x362c37bcU, 0x9d4824f0U,
    0xfdbdd68bU, 0x63258c85U, 0x6726905cU, 0x609500f9U, 0x4de48422U};

__device__ __forceinline__ uint32_t hash_raw(int h, uint32_t x) {  // :13-22
  x ^= x >> 16;
  x *= c_hash_consts[2 * h];
  x ^= x >> 13;
  x *= c_hash_consts[2 * h + 1];
  x ^= x >> 16;
  return x;
}

struct Bloom {
  uint32_t* w;
  uint32_t words;   // power of two
  uint32_t lgw;     // log2(words)
  uint32_t bits;    // words * 32
  int hashes;

  __device__ __forceinline__ void add(uint32_t x) const {  // :8-11, :31-34
    for (int h = 0; h < hashes; ++h) {
      uint32_t pos = hash_raw(h, x) & (bits - 1u);
      atomicOr(&w[pos & (words - 1)], 1u << (pos >> lgw));
    }
  }
  __device__ __forceinline__ bool test(uint32_t x) const {  // :25-28, :37-43
    bool ok = true;
    for (int h = 0; h < hashes; ++h) {
      uint32_t pos = hash_raw(h, x) & (bits - 1u);
      ok &= ((w[pos & (words - 1)] >> (pos >> lgw)) & 1u) != 0;
    }
    return ok;
  }
};

// a > b under the configured comparator (P7)
__device__ __forceinline__ bool key_gt(float ad, int ai, float bd, int bi, bool tuned) {
  if (!tuned) return ad > bd;
  if (ad != bd) return ad > bd;
  return abs(ai) > abs(bi);
}

// One compare-exchange of element e (held by this lane) with partner e^sub
// whose value is (pd, pi).  desc: the pair must end larger-first.
__device__ __forceinline__ void cx(float& md, int& mi, float pd, int pi, bool is_lower,
                                   bool desc, bool tuned) {
  float lo_d = is_lower ? md : pd, hi_d = is_lower ? pd : md;
  int lo_i = is_lower ? mi : pi, hi_i = is_lower ? pi : mi;
  bool swap = desc ? key_gt(hi_d, hi_i, lo_d, lo_i, tuned) : key_gt(lo_d, lo_i, hi_d, hi_i, tuned);
  if (swap) {
    md = pd;
    mi = pi;
  }
}

template <int SUB>
__device__ __forceinline__ float xchg(float v) {
  return __shfl_xor(v, SUB, 64);
}
template <int SUB>
__device__ __forceinline__ int xchg(int v) {
  return __shfl_xor(v, SUB, 64);
}

// Network stage over an R-register array (element e = r*64 + lane) at a
// compile-time distance SUB.  dir(e) gives the direction of e's pair.
template <int R, int SUB, typename DirF>
__device__ __forceinline__ void net_stage(float (&d)[R], int (&id)[R], int lane, bool tuned,
                                          DirF dir) {
  if constexpr (SUB >= 64) {
    constexpr int RS = SUB / 64;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((r & RS) == 0) {
        const int e = r * 64 + lane;
        const bool desc = dir(e);
        float ld = d[r], hd = d[r + RS];
        int li = id[r], hi = id[r + RS];
        bool swap = desc ? key_gt(hd, hi, ld, li, tuned) : key_gt(ld, li, hd, hi, tuned);
        if (swap) {
          d[r] = hd;
          id[r] = hi;
          d[r + RS] = ld;
          id[r + RS] = li;
        }
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = r * 64 + lane;
      float pd = xchg<SUB>(d[r]);
      int pi = xchg<SUB>(id[r]);
      cx(d[r], id[r], pd, pi, (e & SUB) == 0, dir(e), tuned);
    }
  }
}

// Bitonic sort, descending overall, of the first m elements (m a power of
// two, m <= 64*R): graph_search.cuh:888-919.
template <int R, int STEP, int SUB>
__device__ __forceinline__ void sort_sub(float (&d)[R], int (&id)[R], int lane, bool tuned) {
  net_stage<R, SUB>(d, id, lane, tuned, [](int e) { return (e & (2 * STEP)) == 0; });
  if constexpr (SUB > 1) sort_sub<R, STEP, SUB / 2>(d, id, lane, tuned);
}
template <int R, int STEP>
__device__ __forceinline__ void sort_steps(float (&d)[R], int (&id)[R], int m, int lane,
                                           bool tuned) {
  if (STEP <= m / 2) {
    sort_sub<R, STEP, STEP>(d, id, lane, tuned);
    if constexpr (STEP * 2 <= 32 * R) sort_steps<R, STEP * 2>(d, id, m, lane, tuned);
  }
}

// Half-cleaner cascade, every pair ascending, substeps n/2..1:
// graph_search.cuh:933-958.
template <int R, int SUB>
__device__ __forceinline__ void merge_sub(float (&d)[R], int (&id)[R], int n, int lane,
                                          bool tuned) {
  if (SUB <= n / 2) net_stage<R, SUB>(d, id, lane, tuned, [](int) { return false; });
  if constexpr (SUB > 1) merge_sub<R, SUB / 2>(d, id, n, lane, tuned);
}

// LDS layout of one wave's region (in 32-bit words)
struct WaveLds {
  uint32_t* bloom;  // bloom_words
  int* cid;         // 64*MR compacted point ids
  float* cdist;     // 64*MR compacted distances
  float* sbd;       // 64*MR sorted batch distances
  int* sbi;         // 64*MR sorted batch ids
};

template <int MR>
__host__ __device__ constexpr int lds_words_extra() {
  return 4 * 64 * MR;
}

template <int NR, int MR, int J>
struct SearchCore {
  static constexpr int PB = (32 / J) > 0 ? ((32 / J) > 16 ? 16 : (32 / J)) : 1;

  // distances of the valid ids of B (id in [0,np)); others get MAXD.
  // Returns the number of distances computed.
  static __device__ __forceinline__ int batch_dists(const SearchArgs& a, const WaveLds& s,
                                                    const float (&qv)[J], int lane,
                                                    const int (&bi)[MR], float (&bd)[MR],
                                                    int m) {
    const int half = lane >> 5, l32 = lane & 31;
    int total = 0;
    int cix[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      const int slot = r * 64 + lane;
      const bool valid = slot < m && bi[r] >= 0 && (int64_t)bi[r] < a.np;
      const uint64_t bal = __ballot(valid);
      const int ci = total + (int)__builtin_amdgcn_mbcnt_hi(
                                 (uint32_t)(bal >> 32),
                                 __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
      if (valid) s.cid[ci] = bi[r];
      cix[r] = valid ? ci : -1;
      total += __popcll(bal);
    }
    wave_sync();
    for (int p0 = 0; p0 < total; p0 += 2 * PB) {
      float v[PB][J];
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        const int kk = p0 + 2 * pb + half;
        if (kk < total) {
          const float* row = a.base + (int64_t)s.cid[kk] * a.stride;
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const int dim = l32 + 32 * j;
            v[pb][j] = dim < a.d ? row[dim] : 0.0f;
          }
        } else {
#pragma unroll
          for (int j = 0; j < J; ++j) v[pb][j] = 0.0f;
        }
      }
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        const int kk = p0 + 2 * pb + half;
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const float t = v[pb][j] - qv[j];
          acc = __builtin_fmaf(t, t, acc);
        }
        acc += __shfl_down(acc, 16, 32);
        acc += __shfl_down(acc, 8, 32);
        acc += __shfl_down(acc, 4, 32);
        acc += __shfl_down(acc, 2, 32);
        acc += __shfl_down(acc, 1, 32);
        if (l32 == 0 && kk < total) s.cdist[kk] = acc;
      }
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < MR; ++r) bd[r] = cix[r] >= 0 ? s.cdist[cix[r]] : kMaxD;
    wave_sync();
    return total;
  }

  // sort B descending, min-exchange into L, merge L ascending
  static __device__ __forceinline__ void merge_batch(const WaveLds& s, int lane, float (&ld)[NR],
                                                     int (&li)[NR], float (&bd)[MR],
                                                     int (&bi)[MR], int n, int m, bool tuned,
                                                     bool skip_max) {
    sort_steps<MR, 1>(bd, bi, m, lane, tuned);
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      s.sbd[r * 64 + lane] = bd[r];
      s.sbi[r * 64 + lane] = bi[r];
    }
    wave_sync();
    const int c = n < m ? n : m;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int sl = r * 64 + lane;
      if (sl >= n - c && sl < n) {
        const int sb = sl - n + m;
        const float xd = s.sbd[sb];
        const int xi = s.sbi[sb];
        if (!(skip_max && xd == kMaxD) && key_gt(ld[r], li[r], xd, xi, tuned)) {
          ld[r] = xd;
          li[r] = xi;
        }
      }
    }
    wave_sync();
    merge_sub<NR, 32 * NR>(ld, li, n, lane, tuned);
  }
};

// Expansion-flag ballots of the beam (graph_search.cuh:1014-1031).
template <int NR>
__device__ __forceinline__ void flag_ballots(const int (&li)[NR], int lane, int n, int64_t np,
                                             uint64_t (&bf)[NR], uint64_t (&bz)[NR]) {
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int sl = r * 64 + lane;
    const bool in = sl < n;
    bf[r] = __ballot(in && li[r] > 0 && (int64_t)li[r] < np);
    bz[r] = __ballot(in && li[r] == 0);
  }
}

template <int NR>
__device__ __forceinline__ uint32_t scan_chunk(const uint64_t (&bf)[NR], const uint64_t (&bz)[NR],
                                               int chunk, uint32_t& czero) {
  uint32_t vf = 0, vz = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if ((chunk >> 1) == r) {
      vf = (chunk & 1) ? (uint32_t)(bf[r] >> 32) : (uint32_t)bf[r];
      vz = (chunk & 1) ? (uint32_t)(bz[r] >> 32) : (uint32_t)bz[r];
    }
  const uint32_t zm = vz & ~czero;
  czero |= zm;
  return vf | zm;
}

template <int NR, int MR, int J>
__global__ __launch_bounds__(256) void search_kernel(SearchArgs a) {
  using Core = SearchCore<NR, MR, J>;
  extern __shared__ uint32_t smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int region = a.bloom_words_max + lds_words_extra<MR>();
  uint32_t* base_lds = smem + wave * region;
  const int shift = a.shift;
  const int D = 1 << shift;
  const int W = a.width;
  const int m = W << shift;
  const bool tuned = a.mode != 0;
  const int64_t np = a.np;
  const int32_t npi = (int32_t)np;

  for (int64_t q = (int64_t)blockIdx.x * 4 + wave; q < a.nq; q += (int64_t)gridDim.x * 4) {
    // per-query shape (uneven batches carry their own k and beam)
    int n = a.beam, k = a.k;
    if (a.beam_per_q) {
      n = a.beam_per_q[q];
      k = a.k_per_q[q];
    }
    int bwords = a.bloom_words, bhash = a.bloom_hashes;
    if (a.bloom_auto) {  // auto_tune_bloom sizing per beam
      const int lg = 31 - __clz(n);
      bwords = a.auto_words[lg];
      bhash = a.auto_hashes[lg];
    }
    WaveLds s;
    s.bloom = base_lds;
    s.cid = (int*)(base_lds + a.bloom_words_max);
    s.cdist = (float*)(s.cid + 64 * MR);
    s.sbd = s.cdist + 64 * MR;
    s.sbi = (int*)(s.sbd + 64 * MR);
    Bloom bl{s.bloom, (uint32_t)bwords, (uint32_t)(31 - __clz(bwords)), (uint32_t)bwords * 32u,
             bhash};

    float qv[J];
    const float* qp = a.queries + q * a.q_stride;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int dim = (lane & 31) + 32 * j;
      qv[j] = dim < a.d ? qp[dim] : 0.0f;
    }
    for (int i = lane; i < bwords; i += 64) s.bloom[i] = 0u;  // P1
    wave_sync();

    float ld[NR], bd[MR];
    int li[NR], bi[MR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      ld[r] = kMaxD;
      li[r] = npi;
    }
    int hit = 0;
    if (a.algo == 1) {
      hit = a.hits[q];
      if (hit < 0 || hit >= a.n_leaf) hit = 0;
    }
    const int32_t* prow = a.pids + (int64_t)hit * a.cap;
    uint32_t seed = (uint32_t)(q * 64 + lane);  // :66
    int nE = 0, nX = 0;

    // ---- entry seeding, :479-959
    const int nch = (n + m - 1) / m;
    for (int c = 0; c < nch; ++c) {
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        const int i = r * 64 + lane;
        const int idx = c * m + i;
        int id = npi;
        if (i < m && idx < n) {
          if (a.algo == 1) {
            id = idx < a.cap ? prow[idx] : npi;
          } else {  // :50-54
            seed = seed * 1103515245u + 12345u;
            id = (int)(seed % (uint32_t)np);
          }
          bl.add((uint32_t)id);
        }
        bi[r] = id;
      }
      wave_sync();
      nE += Core::batch_dists(a, s, qv, lane, bi, bd, m);
      Core::merge_batch(s, lane, ld, li, bd, bi, n, m, tuned, false);
    }

    // ---- main loop, :964-1532
    int iter = 0, chunk = 0, hash_it = 0;
    bool done = false;
    uint32_t czero = 0;
    uint32_t mask = n >= 32 ? 0xffffffffu : ((1u << n) - 1u);  // P2
    const int nchunks = (n + 31) >> 5;
    while (!done && iter < a.max_iter) {
      ++iter;
#pragma unroll
      for (int r = 0; r < MR; ++r) bi[r] = npi;
      // pick up to W slots (:985-1031)
      int cnt = 0;
      int vrow[kMaxWidth];
      while (cnt < W && mask != 0u) {
        const int b = __builtin_ctz(mask);
        const int sl = chunk * 32 + b;
        int v = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if ((sl >> 6) == r) {
            v = __builtin_amdgcn_readlane(li[r], sl & 63);
            if (lane == (sl & 63)) li[r] = -li[r];
          }
#pragma unroll
        for (int x = 0; x < kMaxWidth; ++x)
          if (x == cnt) vrow[x] = v < 0 ? -v : v;
        mask &= ~(1u << b);
        ++cnt;
        if (cnt == W) break;
        while (mask == 0u && chunk < nchunks) {
          ++chunk;
          uint64_t bf[NR], bz[NR];
          flag_ballots<NR>(li, lane, n, np, bf, bz);
          mask = chunk < nchunks ? scan_chunk<NR>(bf, bz, chunk, czero) : 0u;
        }
      }
      if (cnt == 0) break;
      // graph rows into B; filter test per row against the filter as it was
      // before that row's own adds (P3); rows past np are not read (P4)
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        const int i = r * 64 + lane;
        const int w = i >> shift;
        if (i < m && w < cnt) {
          int vv = 0;
#pragma unroll
          for (int x = 0; x < kMaxWidth; ++x)
            if (x == w) vv = vrow[x];
          bi[r] = (int64_t)vv < np ? a.graph[((int64_t)vv << shift) + (i & (D - 1))] : npi;
        }
      }
#pragma unroll
      for (int w = 0; w < kMaxWidth; ++w) {
        if (w >= cnt) break;
        nX += (int64_t)vrow[w] < np ? 1 : 0;
        bool pass[MR];
#pragma unroll
        for (int r = 0; r < MR; ++r) {
          const int i = r * 64 + lane;
          pass[r] = false;
          if (i < m && (i >> shift) == w) {
            const int t = bi[r];
            pass[r] = !(t == npi || bl.test((uint32_t)t));
          }
        }
        wave_sync();
#pragma unroll
        for (int r = 0; r < MR; ++r) {
          const int i = r * 64 + lane;
          if (i < m && (i >> shift) == w) {
            if (pass[r]) bl.add((uint32_t)bi[r]);
            else bi[r] = npi;
          }
        }
        wave_sync();
      }
      nE += Core::batch_dists(a, s, qv, lane, bi, bd, cnt << shift);
      Core::merge_batch(s, lane, ld, li, bd, bi, n, m, tuned, true);
      // rescan (:1495-1516)
      {
        uint64_t bf[NR], bz[NR];
        flag_ballots<NR>(li, lane, n, np, bf, bz);
        for (chunk = 0; chunk < nchunks; ++chunk) {
          mask = scan_chunk<NR>(bf, bz, chunk, czero);
          if (mask != 0u) break;
          if (chunk == nchunks - 1) done = true;
        }
      }
      if (hash_it == 4) {  // :1518-1531
        for (int i = lane; i < bwords; i += 64) s.bloom[i] = 0u;
        wave_sync();
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int sl = r * 64 + lane;
          const int x = li[r] < 0 ? -li[r] : li[r];
          if (sl < n && (int64_t)x < np) bl.add((uint32_t)x);
        }
        wave_sync();
        hash_it = 0;
      }
      ++hash_it;
    }

    // ---- output (:1955-1963), plus distances
    const int64_t ob = q * (int64_t)a.k_stride;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int sl = r * 64 + lane;
      if (sl < k) {
        a.ids_out[ob + sl] = li[r] < 0 ? -li[r] : li[r];
        if (a.dists_out) a.dists_out[ob + sl] = ld[r];
      }
    }
    if (a.stats_out && lane == 0) {
      a.stats_out[q * 3 + 0] = nE;
      a.stats_out[q * 3 + 1] = nX;
      a.stats_out[q * 3 + 2] = iter;
    }
    wave_sync();
  }
}

template <int NR, int MR, int J>
static hipError_t launch_t(const SearchArgs& a, int grid, hipStream_t st) {
  const size_t lds = (size_t)4 * (a.bloom_words_max + lds_words_extra<MR>()) * 4;
  hipLaunchKernelGGL((search_kernel<NR, MR, J>), dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}

template <int NR, int MR>
static hipError_t launch_j(const SearchArgs& a, int J, int grid, hipStream_t st) {
  switch (J) {
    case 1:
    case 2: return launch_t<NR, MR, 2>(a, grid, st);
    case 3: return launch_t<NR, MR, 3>(a, grid, st);
    case 4: return launch_t<NR, MR, 4>(a, grid, st);
    case 5:
    case 6:
    case 7:
    case 8: return launch_t<NR, MR, 8>(a, grid, st);
    case 30: return launch_t<NR, MR, 30>(a, grid, st);
    default: return hipErrorInvalidValue;
  }
}

static int j_supported(int J) { return (J >= 1 && J <= 8) || J == 30; }

int search_supported(int beam_max, int m, int d) {
  const int J = (d + 31) / 32;
  if (!j_supported(J)) return 0;
  if (beam_max < 1 || beam_max > 512 || m < 1 || m > 128) return 0;
  return 1;
}

hipError_t launch_search(const SearchArgs& a, int beam_max, int grid, hipStream_t st) {
  const int m = a.width << a.shift;
  const int J = (a.d + 31) / 32;
  const int MRv = m <= 64 ? 1 : 2;
  const int NRv = beam_max <= 64 ? 1 : beam_max <= 128 ? 2 : beam_max <= 256 ? 4 : 8;
  if (MRv == 1) {
    switch (NRv) {
      case 1: return launch_j<1, 1>(a, J, grid, st);
      case 2: return launch_j<2, 1>(a, J, grid, st);
      case 4: return launch_j<4, 1>(a, J, grid, st);
      case 8: return launch_j<8, 1>(a, J, grid, st);
    }
  } else {
    switch (NRv) {
      case 1: return launch_j<1, 2>(a, J, grid, st);
      case 2: return launch_j<2, 2>(a, J, grid, st);
      case 4: return launch_j<4, 2>(a, J, grid, st);
      case 8: return launch_j<8, 2>(a, J, grid, st);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace rte
