// Search kernel instantiations for beams up to 128, one wave per query (one
// translation unit per beam register count, so the build compiles them in
// parallel).
#include "search_impl.h"

namespace rte {
template hipError_t launch_nmw<2, 1, 1, false>(const SearchArgs&, int, int, hipStream_t, int*);
template hipError_t launch_nmw<2, 2, 1, false>(const SearchArgs&, int, int, hipStream_t, int*);
}  // namespace rte
