// Learned entry-leaf predictor (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {
struct Mlp;
void mlp_destroy(Mlp* m);
}  // namespace rte
