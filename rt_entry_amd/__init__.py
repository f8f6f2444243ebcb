"""rt_entry for MI355X: entry-leaf selection + greedy beam search on a
proximity graph, as hand-written HIP kernels for gfx950 behind a C ABI
(include/rt_entry.h, librte.so)."""
from .api import (MODE_REF, MODE_TUNED, Context, Params, RteError, build_kdtree, expand_aabbs,
                  hash_parameter, leaf_capacity, make_params)

__all__ = ["Context", "Params", "make_params", "MODE_REF", "MODE_TUNED", "RteError",
           "build_kdtree", "expand_aabbs", "hash_parameter", "leaf_capacity"]
