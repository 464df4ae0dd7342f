// surrogate.cpp — host orchestration of the MLP far-field surrogate (placeholder until the
// network kernels land: the entry points report SIRIUS_HIP_E_UNSUPPORTED).
#include <hip/hip_runtime.h>

#include "../../include/sirius_hip.h"
#include "sirius_internal.hpp"

namespace sirius {

void mlp_release(Handle&) {}

int surrogate_trace(Handle& H, Batch&, const SiriusHipTraceOpts&, DevOut&, SiriusHipStats*) {
    return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "surrogate not built yet");
}

}  // namespace sirius

extern "C" {
size_t sirius_hip_mlp_param_count(void) { return 0; }
int sirius_hip_mlp_init(SiriusHipHandle, uint64_t) { return SIRIUS_HIP_E_UNSUPPORTED; }
int sirius_hip_mlp_get_params(SiriusHipHandle, float*, size_t) { return SIRIUS_HIP_E_UNSUPPORTED; }
int sirius_hip_mlp_set_params(SiriusHipHandle, const float*, size_t) { return SIRIUS_HIP_E_UNSUPPORTED; }
int sirius_hip_mlp_forward(SiriusHipHandle, const float*, float*, int) { return SIRIUS_HIP_E_UNSUPPORTED; }
int sirius_hip_mlp_loss_grad(SiriusHipHandle, const float*, const int32_t*, const float*, int, float*, float*) {
    return SIRIUS_HIP_E_UNSUPPORTED;
}
int sirius_hip_mlp_train_step(SiriusHipHandle, const float*, const int32_t*, const float*, int, float, float*) {
    return SIRIUS_HIP_E_UNSUPPORTED;
}
}
