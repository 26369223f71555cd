// Host launchers of the warm-start MLP kernels (mr_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mr_mlp.h"

namespace mr {

hipError_t mlp_forward(const MRMlpWeightsC& w, const float* x, uint32_t n, float* y_hat, const MRMlpTraceC* trace,
                       hipStream_t st);
uint32_t mlp_loss_scratch_floats(uint32_t n);
hipError_t mlp_loss(const float* y_hat, const float* t, const float* out_scale, uint32_t n, float delta, float* dy,
                    float* loss, float* scratch, hipStream_t st);
hipError_t mlp_backward(const MRMlpWeightsC& w, const MRMlpTraceC& tr, const float* dy, uint32_t n, float* dz2,
                        float* dz1, const MRMlpGradsC& g, hipStream_t st);
hipError_t mlp_adam(float* w, uint16_t* w_bf16, const float* g, float* m1, float* m2, uint32_t count, float lr,
                    float beta1, float beta2, float eps, uint32_t step, hipStream_t st);

} // namespace mr
