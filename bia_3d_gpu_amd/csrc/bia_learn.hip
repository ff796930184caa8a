// bia_learn.hip -- learned-initialiser step (filled in by the MLP work).
#include "bia_runtime.hpp"

namespace bia {
int step_learn_locked(bia_ctx* ctx, int, double, double, int, const double*, int, double*, unsigned long long*) {
  if (!ctx->mlp_ready) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "learned mode needs trained weights");
    return BIA_ERR_STATE;
  }
  return BIA_ERR_STATE;
}
}  // namespace bia
