// frg_mlp.h -- the learned step-size model: 8 features -> 64 -> 64 (tanh) -> log step size.
//
// Parameter block (FP32 master copy, also the C ABI's layout):
//   W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3          (4801 weights)
//   xmu[8], xsd[8], ymu, ysd                                     (input/output normalisation)
// The tracker reads an FP16 copy of the 4801 weights and accumulates in FP32.
//
// Features of a step proposal, taken when a step is accepted (all FP32):
//   0 t                      3 Newton iterations of the corrector   6 log |dx/dt| = -log |V_t|
//   1 log s (this step)      4 log Newton residual                  7 rejections in this step
//   2 log RK error estimate  5 log |Y|
// Target: log of the step that would have met the RK tolerance exactly on the next accepted
// step, s_next * (tol / err_next)^(1/6).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "frg_device.h"

namespace frg {

constexpr int kMlpIn = 8, kMlpH = 64;
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + kMlpH * kMlpIn, kMlpW2 = kMlpB1 + kMlpH, kMlpB2 = kMlpW2 + kMlpH * kMlpH,
              kMlpW3 = kMlpB2 + kMlpH, kMlpB3 = kMlpW3 + kMlpH, kMlpNW = kMlpB3 + 1;  // 4801
constexpr int kMlpXmu = kMlpNW, kMlpXsd = kMlpXmu + kMlpIn, kMlpYmu = kMlpXsd + kMlpIn, kMlpYsd = kMlpYmu + 1,
              kMlpTotal = kMlpYsd + 1;  // 4819
constexpr int kStepRecord = kMlpIn + 1;  // features + target

__device__ __forceinline__ float fast_tanh(float x) {
    // tanh(x) = 1 - 2 / (e^{2x} + 1), saturating for large |x|
    const float e = __expf(2.0f * fminf(fmaxf(x, -15.f), 15.f));
    return 1.0f - 2.0f * __frcp_rn(e + 1.0f);
}

struct StepModel {
    const __half *w;    // kMlpNW FP16 weights
    float xmu[kMlpIn], xsd[kMlpIn], ymu, ysd;
};

// Group-cooperative evaluation (8 lanes, each owning 8 hidden units per layer).  `x` holds the
// raw features, identical in all lanes; `h1buf` is 64 floats of this group's LDS.  Returns the
// predicted log step size in every lane.
__device__ __forceinline__ float mlp_eval_group(int role, const StepModel &M, const float x[kMlpIn], float *h1buf) {
    float xn[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xn[i] = (x[i] - M.xmu[i]) / M.xsd[i];
    const __half *W1 = M.w + kMlpW1, *B1 = M.w + kMlpB1, *W2 = M.w + kMlpW2, *B2 = M.w + kMlpB2, *W3 = M.w + kMlpW3;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int h = 8 * role + u;
        float a = __half2float(B1[h]);
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) a = fmaf(__half2float(W1[h * kMlpIn + i]), xn[i], a);
        h1buf[h] = fast_tanh(a);
    }
    lds_sync();
    float part = 0.f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int h = 8 * role + u;
        float a = __half2float(B2[h]);
#pragma unroll 16
        for (int k = 0; k < kMlpH; ++k) a = fmaf(__half2float(W2[h * kMlpH + k]), h1buf[k], a);
        part = fmaf(__half2float(W3[h]), fast_tanh(a), part);
    }
    lds_sync();
    const float y = grp_sum(part) + __half2float(M.w[kMlpB3]);
    return y * M.ysd + M.ymu;
}

}  // namespace frg
