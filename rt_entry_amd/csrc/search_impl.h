// Greedy beam search on a proximity graph, one wave64 per query (gfx950).
//
// Restates GraphSearchKernel (graph_search.cuh:56-1964) for CDNA4:
//  - one wavefront owns one query; a 256-thread workgroup carries 4
//    independent queries and never uses a workgroup barrier;
//  - the beam L (N slots) and the expansion batch B (M = width * 2^shift
//    slots) live in VGPRs, slot s in lane s % 64 of register s / 64 (or,
//    for L, an LDS queue that the merge loads into that layout); the
//    reference's compare-exchange networks (bitonic sort of B
//    :888-919/:1424-1453, min-exchange :921-931/:1455-1467, half-cleaner
//    merge of L :933-958/:1469-1493) run as cross-lane exchanges (DPP and
//    gfx950 permlane swaps), so every pair sees exactly the comparison the
//    reference makes;
//  - the visited-set filter (:4-43) lives in LDS, one region per wave;
//  - distances follow the reference association: each 32-lane half-wave
//    owns one point, lane l accumulates dims l+32j as an FMA chain, and the
//    lane partials are summed in the shfl_down tree 16,8,4,2,1 (:686-878),
//    here as a transpose-reduce that yields 16 points per half-wave in 38
//    VALU operations with the same pairings;
//  - neighbour ids that pass the filter are compacted before the gathers, so
//    one iteration issues its loads together and only for points that need a
//    distance; a lane reads its half-wave's ids in 16-byte LDS reads and
//    forms each row address with one 32x32->64 multiply-add.
// Parity decisions P1-P7 match the CPU oracle (zeroed filter, all-ones first
// mask, snapshot filter test per row, no row read for ids >= np,
// barrier-separated substeps, REF/TUNED comparators).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "search.h"

namespace rte {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// graph_search.cuh:89-97; h is a compile-time index after unrolling, so the
// multipliers become literals
__device__ __forceinline__ uint32_t hash_const(int i) {
  constexpr uint32_t k[20] = {
      0x924ed183U, 0xd854fc0aU, 0xecf5e3b7U, 0x1bead407U, 0x28a30449U,
      0xbfc4d99fU, 0x715030e2U, 0xffcfb45bU, 0x6e4ce166U, 0xeb53c362U,
      0xa93c4f40U, 0xcecde0a4U, 0x0288592dU, 0x362c37bcU, 0x9d4824f0U,
      0xfdbdd68bU, 0x63258c85U, 0x6726905cU, 0x609500f9U, 0x4de48422U};
  return k[i];
}

__device__ __forceinline__ uint32_t hash_raw(int h, uint32_t x) {  // :13-22
  x ^= x >> 16;
  x *= hash_const(2 * h);
  x ^= x >> 13;
  x *= hash_const(2 * h + 1);
  x ^= x >> 16;
  return x;
}

constexpr int kMaxHashes = 10;
// hash count per comparator mode: the reference's 3 (graph_search.cuh:5) and
// auto_tune_bloom.h's k, which is 7 for every beam from 32 to 512
template <bool TUNED>
constexpr int kHashes = TUNED ? 7 : 3;

// The filter's hash loops are uniform (hashes is a scalar), and they are
// always run by the whole wave: lanes that must not add OR in 0, so no
// uniform loop ever sits inside a divergent region (where the compiler
// would turn its guards into exec masks).
template <int H>
struct Bloom {
  uint32_t* w;
  uint32_t words;  // power of two
  uint32_t lgw;    // log2(words)
  uint32_t bits;   // words * 32

  // bit positions of x, computed once for a test-then-add
  __device__ __forceinline__ void positions(uint32_t x, uint32_t (&pos)[H]) const {
#pragma unroll
    for (int h = 0; h < H; ++h) pos[h] = hash_raw(h, x) & (bits - 1u);
  }
  __device__ __forceinline__ bool test_pos(const uint32_t (&pos)[H]) const {
    bool ok = true;  // :25-28, :37-43
#pragma unroll
    for (int h = 0; h < H; ++h) ok &= ((w[pos[h] & (words - 1)] >> (pos[h] >> lgw)) & 1u) != 0;
    return ok;
  }
  __device__ __forceinline__ void add_pos(const uint32_t (&pos)[H], bool en) const {
#pragma unroll  // :8-11, :31-34
    for (int h = 0; h < H; ++h)
      atomicOr(&w[pos[h] & (words - 1)], en ? 1u << (pos[h] >> lgw) : 0u);
  }
  __device__ __forceinline__ void add(uint32_t x, bool en) const {
    uint32_t pos[H];
    positions(x, pos);
    add_pos(pos, en);
  }
  __device__ __forceinline__ void clear(int lane) const {
    for (int i = 0; i < (int)words; i += 64)
      if (i + lane < (int)words) w[i + lane] = 0u;
  }
};

// a > b under the configured comparator (P7)
template <bool TUNED>
__device__ __forceinline__ bool key_gt(float ad, int ai, float bd, int bi) {
  if constexpr (!TUNED) {
    return ad > bd;
  } else {
    if (ad != bd) return ad > bd;
    return abs(ai) > abs(bi);
  }
}

// Value of lane ^ S (S a power of two below 64) without LDS traffic:
// quad_perm DPP for 1 and 2, row_half_mirror + quad_perm reverse for 4,
// row_ror:8 for 8, permlane16_swap for 16 and permlane32_swap for 32.
template <int S>
__device__ __forceinline__ int xor_lane(int v, int lane) {
  if constexpr (S == 1) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (S == 2) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (S == 4) {
    // half_mirror: i -> 7-i inside each 8 lanes; then reverse each quad
    const int m = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);
    return __builtin_amdgcn_mov_dpp(m, 0x1B, 0xF, 0xF, true);  // quad_perm [3,2,1,0]
  } else if constexpr (S == 8) {
    return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, true);  // row_ror:8
  } else if constexpr (S == 16) {
    // rows r0..r3 of x=y=v: x' = [v0 v0 v2 v2], y' = [v1 v1 v3 v3]
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((lane & 16) ? r[0] : r[1]);
  } else {
    static_assert(S == 32, "xor distance");
    // halves of x=y=v: x' = [lo lo], y' = [hi hi]
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((lane & 32) ? r[0] : r[1]);
  }
}
template <int S>
__device__ __forceinline__ float xor_lane(float v, int lane) {
  return __int_as_float(xor_lane<S>(__float_as_int(v), lane));
}

// Pair directions: a bitonic-sort step (graph_search.cuh:888-919: a pair is
// descending when (e & 2*STEP) == 0) or the merge (all pairs ascending,
// :933-958).
template <int STEP>
struct SortDir {
  static constexpr bool desc(int e) { return (e & (2 * STEP)) == 0; }
};
struct MergeDir {
  static constexpr bool desc(int) { return false; }
};

// Lanes of register r whose element must end a compare-exchange at
// distance SUB holding the larger key: the lower element of a descending
// pair or the upper element of an ascending one.
template <int SUB, class Dir>
constexpr uint64_t want_larger_mask(int r) {
  uint64_t m = 0;
  for (int l = 0; l < 64; ++l) {
    const int e = r * 64 + l;
    if (((e & SUB) == 0) == Dir::desc(e)) m |= 1ull << l;
  }
  return m;
}

// Lane mask of "the partner's key is strictly larger / smaller" under the
// configured comparator (P7).
template <bool TUNED>
__device__ __forceinline__ void key_masks(float md, int mi, float pd, int pi, uint64_t& pgt,
                                          uint64_t& plt) {
  pgt = __builtin_amdgcn_ballot_w64(pd > md);
  plt = __builtin_amdgcn_ballot_w64(pd < md);
  if constexpr (TUNED) {
    const uint64_t eq = __builtin_amdgcn_ballot_w64(pd == md);
    const int ap = abs(pi), am = abs(mi);
    pgt |= eq & __builtin_amdgcn_ballot_w64(ap > am);
    plt |= eq & __builtin_amdgcn_ballot_w64(ap < am);
  }
}

// One register of a cross-lane stage: the lane that must end with the
// larger key takes its partner's value only when the partner is strictly
// larger, the other lane only when it is strictly smaller -- exactly the
// reference's swap on a strict compare.  The direction pattern is a
// compile-time lane mask, so the choice is scalar-unit work.
template <bool TUNED, int SUB, class Dir, int R, int RA, int r>
__device__ __forceinline__ void stage_reg(float (&d)[R], int (&id)[R], int lane) {
  if constexpr (r < RA) {
    constexpr uint64_t W = want_larger_mask<SUB, Dir>(r);
    const float pd = xor_lane<SUB>(d[r], lane);
    const int pi = xor_lane<SUB>(id[r], lane);
    uint64_t pgt, plt;
    key_masks<TUNED>(d[r], id[r], pd, pi, pgt, plt);
    const bool take = __builtin_amdgcn_inverse_ballot_w64((pgt & W) | (plt & ~W));
    d[r] = take ? pd : d[r];
    id[r] = take ? pi : id[r];
    stage_reg<TUNED, SUB, Dir, R, RA, r + 1>(d, id, lane);
  }
}

// Network stage over the first RA registers of an R-register array (element
// e = r*64 + lane) at a compile-time distance SUB.  A network over the first
// N elements pairs them only among themselves, so registers past
// ceil(N / 64) (padding) are left out.
template <bool TUNED, int R, int SUB, class Dir, int RA = R>
__device__ __forceinline__ void net_stage(float (&d)[R], int (&id)[R], int lane) {
  if constexpr (SUB >= 64) {
    constexpr int RS = SUB / 64;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      if ((r & RS) == 0) {
        // the direction depends on r only (2*STEP >= 128 here)
        const bool dsc = Dir::desc(r * 64);
        const float ld = d[r], hd = d[r + RS];
        const int li = id[r], hi = id[r + RS];
        const bool swap = dsc ? key_gt<TUNED>(hd, hi, ld, li) : key_gt<TUNED>(ld, li, hd, hi);
        d[r] = swap ? hd : ld;
        id[r] = swap ? hi : li;
        d[r + RS] = swap ? ld : hd;
        id[r + RS] = swap ? li : hi;
      }
    }
  } else {
    stage_reg<TUNED, SUB, Dir, R, RA, 0>(d, id, lane);
  }
}

// Bitonic sort, descending overall, of the first M_ elements (M_ a power of
// two, M_ <= 64*R): graph_search.cuh:888-919.  Sizes are compile-time so a
// network is straight-line code with no per-substep guards.
// registers holding the first N elements
template <int N, int R>
constexpr int regs_of() {
  return (N + 63) / 64 < 1 ? 1 : ((N + 63) / 64 > R ? R : (N + 63) / 64);
}
template <bool TUNED, int R, int STEP, int SUB, int RA>
__device__ __forceinline__ void sort_sub(float (&d)[R], int (&id)[R], int lane) {
  net_stage<TUNED, R, SUB, SortDir<STEP>, RA>(d, id, lane);
  if constexpr (SUB > 1) sort_sub<TUNED, R, STEP, SUB / 2, RA>(d, id, lane);
}
template <bool TUNED, int R, int M_, int STEP = 1>
__device__ __forceinline__ void sort_net(float (&d)[R], int (&id)[R], int lane) {
  if constexpr (STEP <= M_ / 2) {
    sort_sub<TUNED, R, STEP, STEP, regs_of<M_, R>()>(d, id, lane);
    sort_net<TUNED, R, M_, STEP * 2>(d, id, lane);
  }
}
// runtime m (a uniform power of two) -> the compile-time network
template <bool TUNED, int R, int M_ = 64 * R>
__device__ __forceinline__ void sort_desc(float (&d)[R], int (&id)[R], int m, int lane) {
  if constexpr (M_ >= 2) {
    if (m == M_) sort_net<TUNED, R, M_>(d, id, lane);
    else sort_desc<TUNED, R, M_ / 2>(d, id, m, lane);
  }
}

// Half-cleaner cascade, every pair ascending, substeps N_/2..1:
// graph_search.cuh:933-958.
template <bool TUNED, int R, int SUB, int RA>
__device__ __forceinline__ void merge_net(float (&d)[R], int (&id)[R], int lane) {
  net_stage<TUNED, R, SUB, MergeDir, RA>(d, id, lane);
  if constexpr (SUB > 1) merge_net<TUNED, R, SUB / 2, RA>(d, id, lane);
}
template <bool TUNED, int R, int N_ = 64 * R>
__device__ __forceinline__ void merge_asc(float (&d)[R], int (&id)[R], int n, int lane) {
  if constexpr (N_ >= 2) {
    if (n == N_) merge_net<TUNED, R, N_ / 2, regs_of<N_, R>()>(d, id, lane);
    else merge_asc<TUNED, R, N_ / 2>(d, id, n, lane);
  }
}

// LDS layout of one wave's region (in 32-bit words)
struct WaveLds {
  uint32_t* bloom;  // bloom_words
  int* cid;         // 64*MR compacted point ids
  float* cdist;     // 64*MR compacted distances
  float* sbd;       // 64*MR sorted batch distances
  int* sbi;         // 64*MR sorted batch ids
  float* qd;        // 64*NR beam distances (beam held in LDS)
  int* qi;          // 64*NR beam ids (beam held in LDS)
};

// Expansion-flag ballots of the beam (graph_search.cuh:1014-1031): bf has
// the slots whose id is in (0, np), i.e. not yet expanded (an expanded slot
// holds -id), bz the slots holding id 0, whose sign cannot carry the flag.
template <int NR>
__device__ __forceinline__ void flag_ballots(const int (&li)[NR], int lane, int n, int64_t np,
                                             uint64_t (&bf)[NR], uint64_t (&bz)[NR]) {
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r * 64 >= n) {  // a register past the beam (uniform n): no slots
      bf[r] = 0;
      bz[r] = 0;
      continue;
    }
    // slot r*64 + lane < n, compared as lane < n - r*64 so that no slot
    // index per register stays live in a VGPR
    const bool in = lane < n - r * 64;
    bf[r] = __ballot(in && li[r] > 0 && (int64_t)li[r] < np);
    bz[r] = __ballot(in && li[r] == 0);
  }
}

// The beam L of one query (n slots, n a power of two up to 64*NR).
//
// Register form (LQ false): slot s sits in lane s % 64 of register s / 64
// for the query's whole life.
//
// LDS form (LQ true): the beam is a queue of 64*NR slots in the wave's LDS
// region, slot s at qd[s] / qi[s].  A merge loads the registers that hold
// slots below n into the register form (load), runs the same networks and
// stores them back (store); everything else (the expansion pick, the flag
// ballots, the filter re-add, the output) reads the queue.  The beam then
// holds no VGPRs across the distance gathers, where the kernel's register
// count peaks (beam 512, d = 128: 127 VGPRs instead of 147, 4 waves per
// SIMD instead of 3).  Both forms run the same comparisons in the same
// order.
template <int NR, bool LQ>
struct Beam;

template <int NR>
struct Beam<NR, false> {
  static constexpr bool kInLds = false;
  float d[NR];
  int i[NR];
  __device__ __forceinline__ void init(int, int32_t npi, int) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      d[r] = kMaxD;
      i[r] = npi;
    }
  }
  // key of slot n - 1 (the beam's largest after a merge)
  __device__ __forceinline__ float last(int n) const {
    const int rl = (n - 1) >> 6, ll = (n - 1) & 63;
    int v = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r == rl) v = __builtin_amdgcn_readlane(__float_as_int(d[r]), ll);
    return __int_as_float(v);
  }
  // id of slot sl (uniform), which is marked expanded (-id)
  __device__ __forceinline__ int take(int sl, int lane) {
    int v = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if ((sl >> 6) == r) {
        v = __builtin_amdgcn_readlane(i[r], sl & 63);
        if (lane == (sl & 63)) i[r] = -i[r];
      }
    return v;
  }
  __device__ __forceinline__ void ballots(int lane, int n, int64_t np, uint64_t (&bf)[NR],
                                          uint64_t (&bz)[NR]) const {
    flag_ballots<NR>(i, lane, n, np, bf, bz);
  }
  // slot r*64 + lane
  __device__ __forceinline__ int id_at(int r, int) const { return i[r]; }
  __device__ __forceinline__ float dist_at(int r, int) const { return d[r]; }
};

template <int NR>
struct Beam<NR, true> {
  static constexpr bool kInLds = true;
  float* qd;
  int* qi;
  __device__ __forceinline__ void init(int n, int32_t npi, int lane) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r * 64 >= n) continue;  // registers past the beam are never read
      qd[r * 64 + lane] = kMaxD;
      qi[r * 64 + lane] = npi;
    }
    wave_sync();
  }
  __device__ __forceinline__ float last(int n) const { return qd[n - 1]; }
  __device__ __forceinline__ int take(int sl, int lane) {
    const int v = __builtin_amdgcn_readfirstlane(qi[sl]);
    if (lane == 0) qi[sl] = -v;
    wave_sync();
    return v;
  }
  __device__ __forceinline__ void ballots(int lane, int n, int64_t np, uint64_t (&bf)[NR],
                                          uint64_t (&bz)[NR]) const {
    int x[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) x[r] = r * 64 < n ? qi[r * 64 + lane] : 0;
    flag_ballots<NR>(x, lane, n, np, bf, bz);
  }
  __device__ __forceinline__ int id_at(int r, int lane) const { return qi[r * 64 + lane]; }
  __device__ __forceinline__ float dist_at(int r, int lane) const { return qd[r * 64 + lane]; }
  __device__ __forceinline__ void load(int n, int lane, float (&d)[NR], int (&i)[NR]) const {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const bool in = r * 64 < n;  // uniform
      d[r] = in ? qd[r * 64 + lane] : kMaxD;
      i[r] = in ? qi[r * 64 + lane] : 0;
    }
  }
  __device__ __forceinline__ void store(int n, int lane, const float (&d)[NR],
                                        const int (&i)[NR]) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r * 64 >= n) continue;
      qd[r * 64 + lane] = d[r];
      qi[r * 64 + lane] = i[r];
    }
    wave_sync();
  }
};


template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// One level of the lane-pair sum at distance O inside each 32-lane
// half-wave: lanes with lane bit O clear return lo + the partner's lo, lanes
// with it set return hi + the partner's hi (partner = lane ^ O).
template <int O>
__device__ __forceinline__ float pair_sum(float lo, float hi, int lane) {
  if constexpr (O == 16) {
    // x = lo, y = hi: x' = [x0 y0 x2 y2], y' = [x1 y1 x3 y3] by 16-lane row,
    // so x' + y' is (own + partner) lo in rows 0/2 and hi in rows 1/3
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(lo), __float_as_uint(hi),
                                                    false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else if constexpr (O == 8) {
    const float x = lo + dpp_f<0x108>(lo);  // row_shl:8, lane+8
    const float y = hi + dpp_f<0x118>(hi);  // row_shr:8, lane-8
    return (lane & 8) ? y : x;
  } else if constexpr (O == 4) {
    const float x = lo + dpp_f<0x104>(lo);  // row_shl:4
    const float y = hi + dpp_f<0x114>(hi);  // row_shr:4
    return (lane & 4) ? y : x;
  } else if constexpr (O == 2) {
    const float x = lo + dpp_f<0x4E>(lo);  // quad_perm xor 2
    const float y = hi + dpp_f<0x4E>(hi);
    return (lane & 2) ? y : x;
  } else {
    static_assert(O == 1, "pair distance");
    const float x = lo + dpp_f<0xB1>(lo);  // quad_perm xor 1
    const float y = hi + dpp_f<0xB1>(hi);
    return (lane & 1) ? y : x;
  }
}

// Sum of CNT per-point lane partials over the 32 lanes of each half-wave by
// recursive halving, distances O = 16, 8, 4, 2, 1 in that order: while a
// lane still holds several points it keeps the half selected by lane bit O
// and adds the partner's partials of those points; once it holds one point
// the remaining levels are plain butterfly sums.  Every partial combines
// lanes l and l^O at the level the shfl_down tree 16,8,4,2,1 does
// (graph_search.cuh:873-878), so each total is bit-identical to the
// reference's.  Afterwards the 2^(5 - log2 CNT) consecutive lanes from
// (l32 >> (5 - log2 CNT)) << (5 - log2 CNT) hold point l32 >> (5 - log2 CNT).
template <int CNT, int O>
__device__ __forceinline__ float treduce(const float (&v)[CNT], int lane) {
  if constexpr (O == 0) {
    static_assert(CNT == 1, "more points than lane levels");
    return v[0];
  } else if constexpr (CNT == 1) {
    const float e[1] = {v[0] + xor_lane<O>(v[0], lane)};
    return treduce<1, O / 2>(e, lane);
  } else {
    float nx[CNT / 2];
#pragma unroll
    for (int i = 0; i < CNT / 2; ++i) nx[i] = pair_sum<O>(v[i], v[i + CNT / 2], lane);
    return treduce<CNT / 2, O / 2>(nx, lane);
  }
}

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

template <bool TUNED, int NR, int MR, int J, int WPQ>
struct SearchCore {
  // points per half-wave in one load group: all 16 when the vectors fit in
  // 64 VGPRs, so one iteration's gathers are issued together; with WPQ
  // waves per query each wave takes 1/WPQ of them
  static constexpr int PB0 = (64 / J) >= 16 ? 16 : ((64 / J) > 0 ? (64 / J) : 1);
  static constexpr int PB = PB0 / WPQ > 0 ? PB0 / WPQ : 1;
  static constexpr int SH = 5 - ilog2c(PB);  // a point's lanes: l32 >> SH

  // The PB ids of one load group for this lane's half-wave.  Compacted
  // point kk sits at cid[(kk & 1) * 32*MR + kk / 2], so the points
  // p0 + 2*pb + half of a group are consecutive words and arrive in 16-byte
  // LDS reads.
  static __device__ __forceinline__ void group_ids(const int* c, int (&ids)[PB]) {
    if constexpr (PB >= 4) {
#pragma unroll
      for (int i = 0; i < PB / 4; ++i) {
        const int4 x = reinterpret_cast<const int4*>(c)[i];
        ids[4 * i] = x.x;
        ids[4 * i + 1] = x.y;
        ids[4 * i + 2] = x.z;
        ids[4 * i + 3] = x.w;
      }
    } else if constexpr (PB == 2) {
      const int2 x = reinterpret_cast<const int2*>(c)[0];
      ids[0] = x.x;
      ids[1] = x.y;
    } else {
      ids[0] = c[0];
    }
  }

  // Row values of the group: lane l32 of a half-wave loads dims l32 + 32j
  // of its half's points, from a 64-bit row address made by one 32x32->64
  // multiply-add.  The launcher picks J = ceil(d / 32) for J = 3, 4 and 30,
  // so there only the last dim group can pass d: a lane whose last dim is
  // past d reads its row's first dim instead (in bounds) and takes 0, with
  // no branch.  J = 2 and 8 also serve smaller d (1..32, 129..224): every
  // dim group is masked there.
  static __device__ __forceinline__ void group_rows(const SearchArgs& a, const int (&ids)[PB],
                                                    int l32, float (&v)[PB][J]) {
    const char* bp = reinterpret_cast<const char*>(a.base) + 4 * l32;
    const uint32_t rowb = (uint32_t)a.stride * 4u;
    if constexpr (J == 2 || J == 8) {
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        const float* row = reinterpret_cast<const float*>(
            bp + (uint64_t)(uint32_t)ids[pb] * (uint64_t)rowb);
#pragma unroll
        for (int j = 0; j < J; ++j) v[pb][j] = (l32 + 32 * j) < a.d ? row[32 * j] : 0.0f;
      }
    } else {
      const bool last_in = (l32 + 32 * (J - 1)) < a.d;
      const int lastj = last_in ? 32 * (J - 1) : 0;
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        const float* row = reinterpret_cast<const float*>(
            bp + (uint64_t)(uint32_t)ids[pb] * (uint64_t)rowb);
#pragma unroll
        for (int j = 0; j < J - 1; ++j) v[pb][j] = row[32 * j];
        const float x = row[lastj];
        v[pb][J - 1] = last_in ? x : 0.0f;
      }
    }
  }

  // distances of the valid ids of B (id in [0,np)); others get MAXD.
  // Every wave of the query compacts the same ids (its own copy); wave `sub`
  // computes the load groups sub, sub + WPQ, ... into cdist (shared by the
  // query's waves when WPQ > 1, a workgroup barrier before the reads).
  // Returns the number of distances computed.
  static __device__ __forceinline__ int batch_dists(const SearchArgs& a, const WaveLds& s,
                                                    float* cdist, const float (&qv)[J],
                                                    int lane, int sub, const int (&bi)[MR],
                                                    float (&bd)[MR], int m) {
    constexpr int HALF = 32 * MR;  // cid words per half-wave
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
      if (valid) s.cid[(ci & 1) * HALF + (ci >> 1)] = bi[r];
      cix[r] = valid ? ci : -1;
      total += __popcll(bal);
    }
    // the last group is padded to 2*PB points with row 0 (always a valid
    // row; its distances are discarded)
    const int padded = (total + 2 * PB - 1) & ~(2 * PB - 1);
    if (lane < padded - total) {
      const int kk = total + lane;
      s.cid[(kk & 1) * HALF + (kk >> 1)] = 0;
    }
    wave_sync();
    const int* cbase = s.cid + half * HALF;
    for (int p0 = 2 * PB * sub; p0 < total; p0 += 2 * PB * WPQ) {
      int ids[PB];
      group_ids(cbase + (p0 >> 1), ids);
      float v[PB][J];
      group_rows(a, ids, l32, v);
      float part[PB];
#pragma unroll
      for (int pb = 0; pb < PB; ++pb) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const float t = v[pb][j] - qv[j];
          acc = __builtin_fmaf(t, t, acc);
        }
        part[pb] = acc;
      }
      const float tot = treduce<PB, 16>(part, lane);
      const int kk = p0 + 2 * (l32 >> SH) + half;
      if ((l32 & ((1 << SH) - 1)) == 0 && kk < total) cdist[kk] = tot;
    }
    if constexpr (WPQ > 1) {
      __syncthreads();  // the query's waves are the workgroup
    } else {
      wave_sync();
    }
#pragma unroll
    for (int r = 0; r < MR; ++r) bd[r] = cix[r] >= 0 ? cdist[cix[r]] : kMaxD;
    wave_sync();
    return total;
  }

  // True when no batch element can enter the beam, so the step below would
  // leave L as it is.  The half-cleaner cascade that produced L leaves every
  // comparator relation of all its stages true (its first stage puts each
  // lower element at or below its partner, and comparator networks keep
  // element-wise order, so the recursion keeps it), hence a second pass
  // swaps nothing and L[n-1] is the beam's largest key.  Every element of
  // B not below L[n-1] therefore loses its min-exchange, and with no swap the
  // merge is that second pass.  Distances are non-negative floats, so their
  // bit patterns order as unsigned integers.
  static __device__ __forceinline__ bool batch_cannot_enter(float last, const float (&bd)[MR],
                                                            int lane) {
    int mn = -1;  // 0xffffffff: no valid element
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      const uint32_t u = bd[r] != kMaxD ? __float_as_uint(bd[r]) : 0xffffffffu;
      mn = (int)min((uint32_t)mn, u);
    }
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<1>(mn, lane));
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<2>(mn, lane));
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<4>(mn, lane));
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<8>(mn, lane));
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<16>(mn, lane));
    mn = (int)min((uint32_t)mn, (uint32_t)xor_lane<32>(mn, lane));
    const uint32_t bmin = (uint32_t)__builtin_amdgcn_readfirstlane(mn);
    if (bmin == 0xffffffffu) return true;
    // REF swaps on a strictly larger distance; TUNED also on an equal
    // distance with a larger id, so only a strictly smaller L[n-1] is safe
    return TUNED ? last < __uint_as_float(bmin) : last <= __uint_as_float(bmin);
  }

  // min-exchange of L's last c = min(n, m) slots with the sorted batch's
  // last c (in LDS), then the merge of L
  static __device__ __forceinline__ void exchange_merge(const WaveLds& s, int lane,
                                                        float (&ld)[NR], int (&li)[NR], int n,
                                                        int m, bool skip_max) {
    const int c = n < m ? n : m;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      // registers outside the exchanged slots [n - c, n) are skipped by a
      // scalar branch (n and c are uniform); slot r*64 + lane is compared
      // as lane against scalars, so no slot index stays live in a VGPR
      if ((r + 1) * 64 <= n - c || r * 64 >= n) continue;
      if (lane >= n - c - r * 64 && lane < n - r * 64) {
        const int sb = lane + (r * 64 - n + m);
        const float xd = s.sbd[sb];
        const int xi = s.sbi[sb];
        if (!(skip_max && xd == kMaxD) && key_gt<TUNED>(ld[r], li[r], xd, xi)) {
          ld[r] = xd;
          li[r] = xi;
        }
      }
    }
    wave_sync();
    merge_asc<TUNED, NR>(ld, li, n, lane);
  }

  // sort B descending, min-exchange into L, merge L ascending
  template <class BeamT>
  static __device__ __forceinline__ void merge_batch(const WaveLds& s, int lane, BeamT& L,
                                                     float (&bd)[MR], int (&bi)[MR], int n, int m,
                                                     bool skip_max) {
    if (batch_cannot_enter(L.last(n), bd, lane)) return;
    sort_desc<TUNED, MR>(bd, bi, m, lane);
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      s.sbd[r * 64 + lane] = bd[r];
      s.sbi[r * 64 + lane] = bi[r];
    }
    wave_sync();
    if constexpr (BeamT::kInLds) {
      float ld[NR];
      int li[NR];
      L.load(n, lane, ld, li);
      exchange_merge(s, lane, ld, li, n, m, skip_max);
      L.store(n, lane, ld, li);
    } else {
      exchange_merge(s, lane, L.d, L.i, n, m, skip_max);
    }
  }
};

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

// One scalar fetch from a device-wide counter by a query unit: lane 0 of the
// unit's first wave increments, every wave of the unit gets the value (a
// workgroup barrier and an LDS slot when the unit is several waves).
template <int WPQ>
__device__ __forceinline__ unsigned long long unit_fetch(unsigned long long* ctr, int wave,
                                                         int lane, int64_t* bcast,
                                                         unsigned long long clear_at) {
  unsigned long long i = 0;
  if (wave % WPQ == 0 && lane == 0) {
    i = atomicAdd(ctr, 1ull);
    if (i == clear_at) atomicExch(ctr, 0ull);
    if constexpr (WPQ > 1) *bcast = (int64_t)i;
  }
  if constexpr (WPQ > 1) {
    __syncthreads();
    i = (unsigned long long)*bcast;
    __syncthreads();  // the slot is rewritten only after every wave read it
    return i;
  } else {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)i);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(i >> 32));
    return ((unsigned long long)hi << 32) | lo;
  }
}

// Next query slot of this unit (a wave, or the WPQ waves of a workgroup that
// share one query).  Without a work counter: a grid stride.  With one
// (persistent launch): the first slot of every unit is its global unit id
// (no atomic), later ones come from the counter, which counts past the
// statically assigned prefix and is zero again when the launch ends.  With
// `segments` > 1 the slots are cut into that many contiguous ranges, one per
// group of workgroups blockIdx % segments (the workgroups of one XCD under
// round-robin dispatch), each range with its own counter (reset before the
// launch); a unit whose range is exhausted takes slots from the next ranges.
// Slots map to queries through the order array when there is one.
template <int WPQ>
__device__ __forceinline__ int64_t next_query(const SearchArgs& a, int64_t prev_slot, int wave,
                                              int lane, int64_t* bcast) {
  const int qpb = (blockDim.x >> 6) / WPQ;  // queries per workgroup
  const int64_t units = (int64_t)gridDim.x * qpb;
  const int qw = wave / WPQ;
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(a.work_counter);
  if (!ctr) return prev_slot < 0 ? (int64_t)blockIdx.x * qpb + qw : prev_slot + units;
  if (a.segments <= 1) {
    if (prev_slot < 0) return (int64_t)blockIdx.x * qpb + qw;
    if (units >= a.nq) return a.nq;  // every query was assigned statically
    // a launch makes nq - units increments that get a query and one per
    // unit that does not: nq in all, so the one returning nq - 1 is the
    // last, and it clears the counter for the next launch
    return units + (int64_t)unit_fetch<WPQ>(ctr, wave, lane, bcast,
                                            (unsigned long long)(a.nq - 1));
  }
  const int S = a.segments;
  const int x = (int)(blockIdx.x % (unsigned)S);
  if (prev_slot < 0) {
    const int64_t lu = (int64_t)(blockIdx.x / (unsigned)S) * qpb + qw;
    const int64_t b = a.nq * x / S, e = a.nq * (x + 1) / S;
    if (b + lu < e) return b + lu;
  }
  for (int t = 0; t < S; ++t) {
    const int y = (x + t) % S;
    const int64_t b = a.nq * y / S, e = a.nq * (y + 1) / S;
    // units statically given a slot of range y: one per workgroup of group y
    // (signed: a grid smaller than S leaves groups without workgroups)
    const int64_t uy = (((int64_t)gridDim.x - y + S - 1) / S) * qpb;
    if (uy >= e - b) continue;
    const int64_t i = (int64_t)unit_fetch<WPQ>(ctr + 16 * y, wave, lane, bcast, ~0ull);
    if (b + uy + i < e) return b + uy + i;
  }
  return a.nq;
}

__device__ __forceinline__ int64_t query_of(const SearchArgs& a, int64_t slot) {
  return a.order ? a.order[slot] : slot;
}

// Distance buffer of the next batch: the wave's own, or one of the query's
// two shared buffers, alternating per batch so a wave that runs ahead into
// the next batch never writes the buffer another wave is still reading
// (between two uses of one buffer every wave passes a barrier after its
// reads).
template <int MR, int WPQ>
__device__ __forceinline__ float* dist_buf_t(const WaveLds& s, uint32_t* qshared, int& parity) {
  if constexpr (WPQ == 1) {
    return s.cdist;
  } else {
    float* p = reinterpret_cast<float*>(qshared) + (parity & 1) * 64 * MR;
    parity ^= 1;
    return p;
  }
}
#define dist_buf(s, q, par) dist_buf_t<MR, WPQ>(s, q, par)

// LDS words reserved for a filter of w words: a multiple of four, so the
// compacted ids after it are 16-byte aligned for vector reads
__host__ __device__ constexpr int filter_words(int w) { return (w + 3) & ~3; }

// LDS words of one wave's private region past the filter: compacted ids,
// sorted batch (ids, dists), for one wave per query the distances, and for
// a beam held in LDS its queue of 64*NR (distance, id) slots
template <int MR, int WPQ, int NR, bool LQ>
__host__ __device__ constexpr int lds_words_wave() {
  return (WPQ > 1 ? 3 : 4) * 64 * MR + (LQ ? 2 * 64 * NR : 0);
}
// LDS words a query shares between its waves (WPQ > 1): two distance
// buffers (alternating batches) and the broadcast slot of the work queue
template <int MR, int WPQ>
__host__ __device__ constexpr int lds_words_query() {
  return WPQ > 1 ? 2 * 64 * MR + 4 : 0;
}

template <bool TUNED, int NR, int MR, int J, int WPQ, bool LQ>
__global__ __launch_bounds__(256) void search_kernel(SearchArgs a) {
  using Core = SearchCore<TUNED, NR, MR, J, WPQ>;
  extern __shared__ uint32_t smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = wave % WPQ;  // this wave's share of the query's distances
  const int fwords = filter_words(a.bloom_words_max);
  const int region = fwords + lds_words_wave<MR, WPQ, NR, LQ>();
  uint32_t* base_lds = smem + wave * region;
  // shared by the query's waves (WPQ > 1; one query per workgroup)
  uint32_t* qshared = smem + (blockDim.x >> 6) * region;
  int64_t* bcast = reinterpret_cast<int64_t*>(qshared + 2 * 64 * MR);
  int parity = 0;  // alternates the shared distance buffers
  const int shift = a.shift;
  const int D = 1 << shift;
  const int W = a.width;
  const int m = W << shift;
  const int64_t np = a.np;
  const int32_t npi = (int32_t)np;
  // the uniform shape as register values: with `n = a.beam` inside the
  // loop, the per-query choice below became a select between the argument
  // block's address and the array's, i.e. a generic (flat) load per query
  const int beam_all = __builtin_amdgcn_readfirstlane(a.beam);
  const int k_all = __builtin_amdgcn_readfirstlane(a.k);

  for (int64_t slot = next_query<WPQ>(a, -1, wave, lane, bcast); slot < a.nq;
       slot = next_query<WPQ>(a, slot, wave, lane, bcast)) {
    const int64_t q = query_of(a, slot);
    // per-query shape (uneven batches carry their own k and beam)
    int n = beam_all, k = k_all;
    if (a.beam_per_q) {
      n = a.beam_per_q[q];
      k = a.k_per_q[q];
    }
    // per-query shapes are wave-uniform: keep them in SGPRs so loops over
    // them are scalar branches, not exec-masked ones
    n = __builtin_amdgcn_readfirstlane(n);
    k = __builtin_amdgcn_readfirstlane(k);
    int bwords = a.bloom_words;
    if (a.bloom_auto) {  // auto_tune_bloom sizing per beam
      const int lg = 31 - __clz(n);
      bwords = 1 << (int)((a.auto_lgw >> (4 * lg)) & 15u);
    }
    bwords = __builtin_amdgcn_readfirstlane(bwords);
    WaveLds s;
    s.bloom = base_lds;
    s.cid = (int*)(base_lds + fwords);
    s.sbd = (float*)(s.cid + 64 * MR);
    s.sbi = (int*)(s.sbd + 64 * MR);
    s.cdist = WPQ > 1 ? nullptr : (float*)(s.sbi + 64 * MR);
    s.qd = LQ ? (float*)(base_lds + fwords + lds_words_wave<MR, WPQ, 0, false>())
              : nullptr;
    s.qi = LQ ? (int*)(s.qd + 64 * NR) : nullptr;
    // the hash count is fixed per mode (kHashes), checked on the host
    Bloom<kHashes<TUNED>> bl{s.bloom, (uint32_t)bwords, (uint32_t)(31 - __clz(bwords)),
                             (uint32_t)bwords * 32u};

    float qv[J];
    const float* qp = a.queries + q * a.q_stride;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int dim = (lane & 31) + 32 * j;
      qv[j] = dim < a.d ? qp[dim] : 0.0f;
    }
    bl.clear(lane);  // P1
    wave_sync();

    float bd[MR];
    int bi[MR];
    Beam<NR, LQ> L;
    if constexpr (LQ) {
      L.qd = s.qd;
      L.qi = s.qi;
    }
    L.init(n, npi, lane);
    int hit = 0;
    if (a.algo == 1) {
      hit = __builtin_amdgcn_readfirstlane(a.hits[q]);
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
        const bool in = i < m && idx < n;
        if (in) {
          if (a.algo == 1) {
            id = idx < a.cap ? prow[idx] : npi;
          } else {  // :50-54
            seed = seed * 1103515245u + 12345u;
            id = (int)(seed % (uint32_t)np);
          }
        }
        bl.add((uint32_t)id, in);
        bi[r] = id;
      }
      wave_sync();
      nE += Core::batch_dists(a, s, dist_buf(s, qshared, parity), qv, lane, sub, bi, bd, m);
      Core::merge_batch(s, lane, L, bd, bi, n, m, false);
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
#pragma unroll
      for (int x = 0; x < kMaxWidth; ++x) vrow[x] = 0;
      while (cnt < W && mask != 0u) {
        const int b = __builtin_ctz(mask);
        const int sl = chunk * 32 + b;
        const int v = L.take(sl, lane);
#pragma unroll
        for (int x = 0; x < kMaxWidth; ++x)
          if (x == cnt) vrow[x] = v < 0 ? -v : v;
        mask &= ~(1u << b);
        ++cnt;
        if (cnt == W) break;
        while (mask == 0u && chunk < nchunks) {
          ++chunk;
          uint64_t bf[NR], bz[NR];
          L.ballots(lane, n, np, bf, bz);
          mask = chunk < nchunks ? scan_chunk<NR>(bf, bz, chunk, czero) : 0u;
        }
      }
      if (cnt == 0) break;
      // graph rows into B (rows past np are not read, P4)
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
      // filter test per row against the filter as it was before that row's
      // own adds (P3), then the adds
#pragma unroll
      for (int w = 0; w < kMaxWidth; ++w) {
        if (w >= cnt) break;
        nX += (int64_t)vrow[w] < np ? 1 : 0;
        uint32_t pos[MR][kHashes<TUNED>];
        bool pass[MR];
#pragma unroll
        for (int r = 0; r < MR; ++r) {
          const int i = r * 64 + lane;
          const bool mine = i < m && (i >> shift) == w;
          const int t = bi[r];
          bl.positions((uint32_t)t, pos[r]);
          const bool seen = bl.test_pos(pos[r]);
          pass[r] = mine && t != npi && !seen;
        }
        wave_sync();
#pragma unroll
        for (int r = 0; r < MR; ++r) {
          const int i = r * 64 + lane;
          const bool mine = i < m && (i >> shift) == w;
          bl.add_pos(pos[r], pass[r]);
          bi[r] = (mine && !pass[r]) ? npi : bi[r];
        }
        wave_sync();
      }
      nE += Core::batch_dists(a, s, dist_buf(s, qshared, parity), qv, lane, sub, bi, bd,
                              cnt << shift);
      Core::merge_batch(s, lane, L, bd, bi, n, m, true);
      // rescan (:1495-1516)
      {
        uint64_t bf[NR], bz[NR];
        L.ballots(lane, n, np, bf, bz);
        for (chunk = 0; chunk < nchunks; ++chunk) {
          mask = scan_chunk<NR>(bf, bz, chunk, czero);
          if (mask != 0u) break;
          if (chunk == nchunks - 1) done = true;
        }
      }
      if (hash_it == 4) {  // :1518-1531
        bl.clear(lane);
        wave_sync();
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          if (r * 64 >= n) continue;  // past the beam (uniform)
          const int id = L.id_at(r, lane);
          const int x = id < 0 ? -id : id;
          bl.add((uint32_t)x, lane < n - r * 64 && (int64_t)x < np);
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
      if (r * 64 >= k) continue;  // past the results (uniform k)
      const int sl = r * 64 + lane;
      if (sub == 0 && lane < k - r * 64) {
        const int id = L.id_at(r, lane);
        a.ids_out[ob + sl] = id < 0 ? -id : id;
        if (a.dists_out) a.dists_out[ob + sl] = L.dist_at(r, lane);
      }
    }
    if (a.stats_out && sub == 0 && lane == 0) {
      a.stats_out[q * 3 + 0] = nE;
      a.stats_out[q * 3 + 1] = nX;
      a.stats_out[q * 3 + 2] = iter;
    }
    wave_sync();
  }
}

template <int MR, int WPQ, int NR, bool LQ>
size_t search_lds_bytes(int bloom_words_max, int wpb) {
  return ((size_t)wpb * (filter_words(bloom_words_max) + lds_words_wave<MR, WPQ, NR, LQ>()) +
          lds_words_query<MR, WPQ>()) * 4;
}

template <bool TUNED, int NR, int MR, int J, int WPQ, bool LQ>
hipError_t launch_t(const SearchArgs& a, int grid, int wpb, hipStream_t st, int* occ) {
  const size_t lds = search_lds_bytes<MR, WPQ, NR, LQ>(a.bloom_words_max, wpb);
  if (occ) {  // resident workgroups per CU for this shape
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        occ, search_kernel<TUNED, NR, MR, J, WPQ, LQ>, 64 * wpb, lds);
  }
  hipLaunchKernelGGL((search_kernel<TUNED, NR, MR, J, WPQ, LQ>), dim3(grid), dim3(64 * wpb), lds,
                     st, a);
  return hipGetLastError();
}

template <int NR, int MR, int WPQ, bool LQ>
hipError_t launch_nmw(const SearchArgs& a, int grid, int wpb, hipStream_t st, int* occ) {
  const int J = (a.d + 31) / 32;
  const bool tuned = a.mode != 0;
#define RTE_J_CASES(T)                                                        \
  switch (J) {                                                                \
    case 1:                                                                   \
    case 2: return launch_t<T, NR, MR, 2, WPQ, LQ>(a, grid, wpb, st, occ);    \
    case 3: return launch_t<T, NR, MR, 3, WPQ, LQ>(a, grid, wpb, st, occ);    \
    case 4: return launch_t<T, NR, MR, 4, WPQ, LQ>(a, grid, wpb, st, occ);    \
    case 5:                                                                   \
    case 6:                                                                   \
    case 7:                                                                   \
    case 8: return launch_t<T, NR, MR, 8, WPQ, LQ>(a, grid, wpb, st, occ);    \
    case 30: return launch_t<T, NR, MR, 30, WPQ, LQ>(a, grid, wpb, st, occ);  \
    default: return hipErrorInvalidValue;                                     \
  }
  if (tuned) {
    RTE_J_CASES(true)
  } else {
    RTE_J_CASES(false)
  }
#undef RTE_J_CASES
}

}  // namespace rte
