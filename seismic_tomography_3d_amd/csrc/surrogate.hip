// surrogate.hip -- learned warm start (filled in with the network kernels).
#include "st_common.hpp"

namespace st {
bool surrogate_warm_fill(st_ctx *, const float *, int, int, float, float, bool, float *) {
    return false;
}
float surrogate_cold_radius(st_ctx *ctx) { return 2.0f * 8 * ctx->dx; }
void surrogate_release(st_ctx *) {}
}  // namespace st
