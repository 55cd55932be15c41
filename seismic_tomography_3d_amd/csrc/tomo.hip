// tomo.hip -- ray tracing, CGLS and model update (filled in next).
#include "st_common.hpp"

namespace st {
void tomo_release(st_ctx *) {}
}  // namespace st
