// mi_pending.hip -- entry points whose kernels live in later translation units; each reports
// MI_ERR_STATE until its implementation is linked in.
#include "mi_internal.h"

extern "C" {

void mi_flow_release(mi_ctx *) {}

int mi_flow_integrate_internal(mi_ctx *ctx, const mi_run *, mi_sigma_est *)
{
    ctx->err = "flow sampler not available";
    return MI_ERR_STATE;
}

#define PENDING(name, ...)                                                                          \
    int name(mi_ctx *ctx, __VA_ARGS__)                                                              \
    {                                                                                               \
        if (ctx) ctx->err = #name " not available";                                                 \
        return MI_ERR_STATE;                                                                        \
    }

PENDING(mi_adjoint, const mi_run *, mi_sigma_grad *)
PENDING(mi_flow_init, uint64_t)
PENDING(mi_flow_train, const mi_run *, int, uint64_t, double, mi_flow_info *)
PENDING(mi_flow_forward, uint64_t, const float *, float *, float *)
PENDING(mi_flow_inverse, uint64_t, const float *, float *, float *)
PENDING(mi_flow_backward, uint64_t, const float *, const float *, float *)
PENDING(mi_flow_get_params, float *)
PENDING(mi_flow_set_params, const float *)

}  // extern "C"
