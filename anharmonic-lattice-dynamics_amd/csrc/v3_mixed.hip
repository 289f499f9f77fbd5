// Mixed-precision V3 path (FP32 MFMA phase GEMM with hi/lo split).
#include "alt3_impl.h"

namespace alt3 {

int v3_mixed_run(alt3_ctx *ctx, int ksub, int *d_err)
{
    (void)ctx;
    (void)ksub;
    (void)d_err;
    return fail(ALT3_E_ARG, "ALT3_V3_MIXED is not available in this build");
}

}  // namespace alt3
