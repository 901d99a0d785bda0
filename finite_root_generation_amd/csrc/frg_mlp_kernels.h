// frg_mlp_kernels.h -- arguments and launchers of the step-model training kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frg_mlp.h"

namespace frg {

struct MlpGradArgs {
    const float *w;        // kMlpNW FP32 weights
    const float *samples;  // n_samples x kStepRecord (features, target)
    const float *norm;     // xmu[8], xsd[8], ymu, ysd
    const int64_t *indices;  // batch sample rows, or null: Philox draws (seed, step)
    int64_t n_samples, batch;
    uint64_t seed, step;
    float *grad;           // kMlpNW accumulator
    float *loss;           // accumulates the minibatch loss
};

struct MlpAdamArgs {
    float *w, *m, *v, *grad;
    __half *w16;
    float lr, b1, b2, eps, b1t, b2t;  // b1t = b1^step, b2t = b2^step
};

hipError_t launch_mlp_grad(const MlpGradArgs &a, hipStream_t s);
hipError_t launch_mlp_adam(const MlpAdamArgs &a, hipStream_t s);
hipError_t launch_mlp_to_half(const float *w, __half *w16, hipStream_t s);
hipError_t launch_mlp_eval(const StepModel &m, const float *x, int64_t n, float *y, hipStream_t s);

}  // namespace frg
