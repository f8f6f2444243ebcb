// Dispatch of the beam-search kernel (search_impl.h) by beam and batch size.
#include <hip/hip_runtime.h>

#include "search.h"

namespace rte {

static int j_supported(int J) { return (J >= 1 && J <= 8) || J == 30; }

int search_supported(int beam_max, int m, int d) {
  const int J = (d + 31) / 32;
  if (!j_supported(J)) return 0;
  if (beam_max < 1 || beam_max > 512 || m < 1 || m > 128) return 0;
  return 1;
}

static int nr_of(int beam_max) {
  return beam_max <= 64 ? 1 : beam_max <= 128 ? 2 : beam_max <= 256 ? 4 : 8;
}

template <int NR, int MR>
static hipError_t by_wpq(const SearchArgs& a, int grid, int wpb, hipStream_t st, int* occ) {
  if (a.queue == 1) return launch_nmw<NR, MR, 1, true>(a, grid, wpb, st, occ);
  switch (a.wpq) {
    case 0:
    case 1: return launch_nmw<NR, MR, 1, false>(a, grid, wpb, st, occ);
    case 2: return launch_nmw<NR, MR, 2, false>(a, grid, wpb, st, occ);
    case 4: return launch_nmw<NR, MR, 4, false>(a, grid, wpb, st, occ);
  }
  return hipErrorInvalidValue;
}

static hipError_t dispatch(const SearchArgs& a, int beam_max, int grid, int wpb, hipStream_t st,
                           int* occ) {
  const int m = a.width << a.shift;
  const int MRv = m <= 64 ? 1 : 2;
  switch (nr_of(beam_max) * 4 + MRv) {
    case 1 * 4 + 1: return by_wpq<1, 1>(a, grid, wpb, st, occ);
    case 1 * 4 + 2: return by_wpq<1, 2>(a, grid, wpb, st, occ);
    case 2 * 4 + 1: return by_wpq<2, 1>(a, grid, wpb, st, occ);
    case 2 * 4 + 2: return by_wpq<2, 2>(a, grid, wpb, st, occ);
    case 4 * 4 + 1: return by_wpq<4, 1>(a, grid, wpb, st, occ);
    case 4 * 4 + 2: return by_wpq<4, 2>(a, grid, wpb, st, occ);
    case 8 * 4 + 1: return by_wpq<8, 1>(a, grid, wpb, st, occ);
    case 8 * 4 + 2: return by_wpq<8, 2>(a, grid, wpb, st, occ);
  }
  return hipErrorInvalidValue;
}

static bool shape_ok(const SearchArgs& a, int wpb) {
  if (a.queue != 0 && a.queue != 1) return false;
  if (a.queue == 1 && a.wpq > 1) return false;  // the LDS queue: one wave per query
  if (a.wpq > 1) return (a.wpq == 2 || a.wpq == 4) && wpb == a.wpq;
  return a.wpq >= 0 && wpb >= 1 && wpb <= 4;
}

hipError_t launch_search(const SearchArgs& a, int beam_max, int grid, int wpb, hipStream_t st) {
  if (!shape_ok(a, wpb) || grid < 1) return hipErrorInvalidValue;
  return dispatch(a, beam_max, grid, wpb, st, nullptr);
}

hipError_t search_occupancy(const SearchArgs& a, int beam_max, int wpb, int* blocks_per_cu) {
  if (!shape_ok(a, wpb)) return hipErrorInvalidValue;
  return dispatch(a, beam_max, 1, wpb, nullptr, blocks_per_cu);
}

}  // namespace rte
