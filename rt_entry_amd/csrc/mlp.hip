// Learned entry-leaf predictor: placeholder until the MLP kernels land.
#include "ctx.h"
#include "mlp.h"

namespace rte {
struct Mlp {};
void mlp_destroy(Mlp* m) { delete m; }
}  // namespace rte

rte_ctx::~rte_ctx() { rte::mlp_destroy(mlp); }

extern "C" {
int rte_mlp_create(rte_ctx* c, int, int, int, int, uint64_t) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
int rte_mlp_train_step(rte_ctx* c, const float*, const int32_t*, int, float, float, float*) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
int rte_mlp_predict(rte_ctx* c, const float*, int64_t, const int32_t*, float, int32_t*, float*) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
int rte_mlp_get(rte_ctx* c, float*, float*, float*, float*, float*, float*) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
int rte_mlp_set(rte_ctx* c, const float*, const float*, const float*, const float*, const float*,
                const float*) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
int rte_mlp_grad(rte_ctx* c, const float*, const int32_t*, int, float*, float*, float*, float*,
                 float*, float*, float*) {
  return rte_fail(c, RTE_EUNSUP, "predictor not built");
}
}
