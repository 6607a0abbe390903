// Mixed-precision operator GEMMs (FP32 / BF16 MFMA with FP32 accumulation).
#include "ngw_kernels.h"

namespace ngw {

int lowp_prepare(ngw_context* ctx)
{
    (void)ctx;
    return fail(NGW_ERR_STATE, "mixed precision not available in this build");
}

int operator_lowp(ngw_context* ctx, const OperatorArgs& a)
{
    (void)ctx;
    (void)a;
    return fail(NGW_ERR_STATE, "mixed precision not available in this build");
}

}  // namespace ngw
