// unet.hip -- learned preconditioner for flexible CG (placeholder until the
// network kernels land): reports that no trained network is available.
#include "ctx_internal.h"

namespace shk {

struct UnetPrecond {};

void unet_destroy(UnetPrecond *u) { delete u; }

int unet_solve(shk_ctx *c, shk_result *res) {
    (void)res;
    return c->fail(SHK_BAD_ARG, "Preconditioner=unet: no trained network is available in this build");
}

}  // namespace shk
