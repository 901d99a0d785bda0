// frg_mlp.h -- the learned step-size model: 8 features -> 64 -> 64 (tanh) -> log step size.
//
// Parameter block (FP32 master copy, also the C ABI's layout):
//   W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3          (4801 weights)
//   xmu[8], xsd[8], ymu, ysd                                     (input/output normalisation)
// The tracker reads an FP16 copy of the 4801 weights and accumulates in FP32; the first hidden
// layer's activations (tanh, in [-1, 1]) are rounded to FP16 so that the second layer runs as
// FP16 pair dot products with FP32 accumulation (v_dot2_f32_f16).
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

// acc + a.lo * b.lo + a.hi * b.hi for two FP16 pairs packed in 32-bit words (FP32 accumulation).
__device__ __forceinline__ float dot2_f16(uint32_t a, uint32_t b, float acc) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, a), __builtin_bit_cast(h2, b), acc, false);
}

__device__ __forceinline__ void unpack8(uint4 q, float out[8]) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const __half2 h = *reinterpret_cast<const __half2 *>(&w[i]);
        out[2 * i] = __low2float(h);
        out[2 * i + 1] = __high2float(h);
    }
}

// One lane's share of the model: hidden unit l (the lane index) of both layers, i.e. row l of W1
// and of W2 and the unit's b1, b2 and w3.
struct MlpLane {
    uint4 w1;
    uint4 w2[8];
    float b1, b2, w3;
};
// Loads this lane's weights (FP16, from L2).
__device__ __forceinline__ MlpLane mlp_lane_weights(const StepModel &M) {
    const int l = __lane_id();
    MlpLane w;
    w.w1 = reinterpret_cast<const uint4 *>(M.w + kMlpW1 + kMlpIn * l)[0];
    const uint4 *w2 = reinterpret_cast<const uint4 *>(M.w + kMlpW2 + kMlpH * l);
#pragma unroll
    for (int i = 0; i < 8; ++i) w.w2[i] = w2[i];
    w.b1 = __half2float(M.w[kMlpB1 + l]);
    w.b2 = __half2float(M.w[kMlpB2 + l]);
    w.w3 = __half2float(M.w[kMlpW3 + l]);
    return w;
}

// Wave-wide evaluation for the groups of one wave that have a pending proposal; W = 8 or 16
// lanes per group.  Lane l owns hidden unit l of both layers, so its weights are one row of W1
// and one row of W2 (`lw`, from mlp_lane_weights).  Must be called by all 64 lanes with a
// wave-uniform `lanes` mask whose set bits are lanes W*g of the pending groups g.  Group g has
// written its normalised features to area0[g * stride + 0..7] (floats in LDS); the same 128-float
// area then holds that group's first hidden layer (64 FP16 values at float offset +64) and the 64
// output-layer terms (at +0).  `gw` is this lane's group in the wave, `role` its lane in the
// group.  Returns the predicted log step size in the lanes of each pending group.
template <int W = 8>
__device__ __forceinline__ float mlp_eval_wave(const StepModel &M, const MlpLane &lw, float *area0, int stride,
                                               uint64_t lanes, int gw, int role) {
    constexpr int kShift = W == 16 ? 4 : 3;
    const int l = __lane_id();
    float a1[8];
    unpack8(lw.w1, a1);
    lds_sync();
    // layer 1 for every pending group
    for (uint64_t m = lanes; m; m &= m - 1) {
        float *area = area0 + (int)(__builtin_ctzll(m) >> kShift) * stride;
        const float4 x0 = reinterpret_cast<const float4 *>(area)[0];
        const float4 x1 = reinterpret_cast<const float4 *>(area)[1];
        float a = lw.b1;
        a = fmaf(a1[0], x0.x, a);
        a = fmaf(a1[1], x0.y, a);
        a = fmaf(a1[2], x0.z, a);
        a = fmaf(a1[3], x0.w, a);
        a = fmaf(a1[4], x1.x, a);
        a = fmaf(a1[5], x1.y, a);
        a = fmaf(a1[6], x1.z, a);
        a = fmaf(a1[7], x1.w, a);
        reinterpret_cast<__half *>(area + kMlpH)[l] = __float2half(fast_tanh(a));
    }
    lds_sync();
    // layer 2 and the output-layer terms w3[l] tanh(.) for every pending group; four partial
    // sums of 16 products each, so that the dependent chain is 8 dot products long, not 32
    for (uint64_t m = lanes; m; m &= m - 1) {
        float *area = area0 + (int)(__builtin_ctzll(m) >> kShift) * stride;
        const uint4 *hv = reinterpret_cast<const uint4 *>(area + kMlpH);
        float acc[4] = {lw.b2, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint4 wq = lw.w2[i], hq = hv[i];
            acc[0] = dot2_f16(wq.x, hq.x, acc[0]);
            acc[1] = dot2_f16(wq.y, hq.y, acc[1]);
            acc[2] = dot2_f16(wq.z, hq.z, acc[2]);
            acc[3] = dot2_f16(wq.w, hq.w, acc[3]);
        }
        area[l] = lw.w3 * fast_tanh((acc[0] + acc[1]) + (acc[2] + acc[3]));
    }
    lds_sync();
    // each group sums its own 64 terms: 64/W per lane, then over the group
    const float4 *tv = reinterpret_cast<const float4 *>(area0 + gw * stride + (64 / W) * role);
    const float4 t0 = tv[0];
    float part;
    if (W == 8) {
        const float4 t1 = tv[1];
        part = ((t0.x + t0.y) + (t0.z + t0.w)) + ((t1.x + t1.y) + (t1.z + t1.w));
    } else {
        part = (t0.x + t0.y) + (t0.z + t0.w);
    }
    part = grp_sum<W>(part);
    lds_sync();
    const float y = part + __half2float(M.w[kMlpB3]);
    return y * M.ysd + M.ymu;
}

}  // namespace frg
