// Launch interface of the beam-search kernel (search.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {

constexpr float kMaxD = 536870912.0f;  // (float)0x1fffffff, structure_on_device.cuh:6
constexpr int kMaxWidth = 4;           // search_width rows per iteration

struct SearchArgs {
  const float* base;
  int64_t np;
  int64_t stride;
  int d;
  const int32_t* graph;
  int shift;  // log2 of the graph row stride
  const float* queries;
  int64_t q_stride;
  int64_t nq;
  const int32_t* pids;  // n_leaf x cap
  int cap;
  int n_leaf;
  const int32_t* hits;
  int32_t* ids_out;
  float* dists_out;
  int32_t* stats_out;
  int k;
  int beam;
  int k_stride;
  int width;
  int max_iter;
  int mode;
  int algo;
  int bloom_words;
  int bloom_hashes;
  int bloom_words_max;  // LDS words reserved per wave for the filter
  int bloom_auto;       // per-query filter sizing from auto_lgw
  // log2 of the filter words for beam 2^i in bits [4i, 4i + 4) (filter
  // sizes are powers of two; one 64-bit field instead of a table, which the
  // kernel would hold in scalar registers for its whole life)
  uint64_t auto_lgw;
  const int32_t* k_per_q;     // uneven batches (or null)
  const int32_t* beam_per_q;  // uneven batches (or null)
  int64_t* work_counter;      // persistent launch: shared query counter (or null)
  const int32_t* order;       // persistent launch: query order (or null)
  int segments;               // persistent launch: slot ranges with their own
                              // counters (work_counter + 16 * range), 0/1: one
  int wpq;                    // waves per query (1, 2, 4); > 1: one query per
                              // workgroup of wpq waves
  int queue;                  // where the beam lives: 0 VGPRs, 1 an LDS queue
                              // of 64 * ceil(beam bound / 64) slots per wave
                              // (one wave per query)
};

int search_supported(int beam_max, int m, int d);
// wpb: waves per workgroup, 1..4 (one query each), or a.wpq (> 1: one query
// of a.wpq waves per workgroup)
hipError_t launch_search(const SearchArgs& a, int beam_max, int grid, int wpb, hipStream_t st);
// resident workgroups per CU for this launch shape
hipError_t search_occupancy(const SearchArgs& a, int beam_max, int wpb, int* blocks_per_cu);

// launchers per beam register count, batch register count, waves per query
// and beam storage (translation units search_nr*.hip); with occ non-null
// they report occupancy instead of launching
template <int NR, int MR, int WPQ, bool LQ>
hipError_t launch_nmw(const SearchArgs& a, int grid, int wpb, hipStream_t st, int* occ);

}  // namespace rte
