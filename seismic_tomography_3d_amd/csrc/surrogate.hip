// surrogate.hip -- learned warm start (filled in with the network kernels).
#include "st_common.hpp"

namespace st {
bool surrogate_warm_fill(st_ctx *, const float *, int, int, float) { return false; }
int surrogate_judge(st_ctx *, const float *, int, int, float, int *) { return 0; }
void surrogate_release(st_ctx *) {}
}  // namespace st
