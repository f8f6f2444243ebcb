// Search kernel instantiations for beams up to 64 held in an LDS queue, one
// wave per query.
#include "search_impl.h"

namespace rte {
template hipError_t launch_nmw<1, 1, 1, true>(const SearchArgs&, int, int, hipStream_t, int*);
template hipError_t launch_nmw<1, 2, 1, true>(const SearchArgs&, int, int, hipStream_t, int*);
}  // namespace rte
