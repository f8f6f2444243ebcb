// Search kernel instantiations for beams up to 128 with two or four waves
// per query.
#include "search_impl.h"

namespace rte {
template hipError_t launch_nmw<2, 1, 2, false>(const SearchArgs&, int, int, hipStream_t, int*);
template hipError_t launch_nmw<2, 2, 2, false>(const SearchArgs&, int, int, hipStream_t, int*);
template hipError_t launch_nmw<2, 1, 4, false>(const SearchArgs&, int, int, hipStream_t, int*);
template hipError_t launch_nmw<2, 2, 4, false>(const SearchArgs&, int, int, hipStream_t, int*);
}  // namespace rte
