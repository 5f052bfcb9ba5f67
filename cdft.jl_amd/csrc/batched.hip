// batched.hip -- persistent batched solver (placeholder until the kernel lands).
#pragma once
#include "common.h"

namespace cdft {
inline int batched_prepare_mixed(cdft_sys *) { return CDFT_OK; }
}  // namespace cdft

extern "C" int cdft_converge_batch(cdft_ctx *ctx, int, cdft_sys *const *, double *const *, const cdft_solver_opts *,
                                   cdft_result *)
{
    if (ctx) ctx->err = "batched path not built yet";
    return CDFT_ERR_UNSUPPORTED;
}
extern "C" int cdft_converge_batch_device(cdft_ctx *ctx, int, cdft_sys *const *, double *const *,
                                          const cdft_solver_opts *, cdft_result *)
{
    if (ctx) ctx->err = "batched path not built yet";
    return CDFT_ERR_UNSUPPORTED;
}
