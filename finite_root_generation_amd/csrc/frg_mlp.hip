// frg_mlp.hip -- training of the learned step-size model on the GPU.
//
// Loss: mean over a minibatch of (y(x) - t)^2 / 2 in normalised units (x and t standardised by
// the stored means and deviations).  Gradient kernel: each block takes 32 samples, keeps the
// 4801 weights and the per-sample activations in LDS, runs the forward pass, back-propagates,
// reduces the weight gradients over its samples and adds them to a global accumulator.  Adam
// kernel: one thread per weight, also writes the FP16 copy the tracker reads and clears the
// accumulator.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frg_device.h"
#include "frg_mlp.h"
#include "frg_mlp_kernels.h"

namespace frg {

namespace {
constexpr int kSB = 32;  // samples per block (LDS: weights 19 KB + activations 33 KB)
constexpr int kThreads = 256;

__device__ __forceinline__ float tanh_exact(float x) { return tanhf(x); }
}  // namespace

__global__ __launch_bounds__(kThreads) void mlp_grad_kernel(MlpGradArgs A) {
    __shared__ float W[kMlpNW];
    __shared__ float xn[kSB][kMlpIn];
    __shared__ float h1[kSB][kMlpH];
    __shared__ float h2[kSB][kMlpH];
    __shared__ float d2[kSB][kMlpH];
    __shared__ float d1[kSB][kMlpH];
    __shared__ float dy[kSB];
    const int tid = threadIdx.x;
    for (int p = tid; p < kMlpNW; p += kThreads) W[p] = A.w[p];
    // samples of this block
    const int64_t base = (int64_t)blockIdx.x * kSB;
    if (tid < kSB) {
        int64_t s = base + tid;
        bool valid = s < A.batch;
        int64_t row = 0;
        if (valid) {
            if (A.indices) {
                row = A.indices[s];
            } else {
                double a, b;
                uniform_pair((uint64_t)s, A.step, A.seed, a, b);
                row = (int64_t)(a * (double)A.n_samples);
                if (row >= A.n_samples) row = A.n_samples - 1;
            }
        }
        const float *rec = A.samples + row * kStepRecord;
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) xn[tid][i] = valid ? (rec[i] - A.norm[i]) / A.norm[kMlpIn + i] : 0.f;
        dy[tid] = valid ? (rec[kMlpIn] - A.norm[2 * kMlpIn]) / A.norm[2 * kMlpIn + 1] : 0.f;  // target for now
        if (!valid) dy[tid] = __int_as_float(0x7fc00000);                                    // NaN marks padding
    }
    __syncthreads();
    // layer 1
    for (int q = tid; q < kSB * kMlpH; q += kThreads) {
        const int s = q / kMlpH, h = q % kMlpH;
        float a = W[kMlpB1 + h];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) a = fmaf(W[kMlpW1 + h * kMlpIn + i], xn[s][i], a);
        h1[s][h] = tanh_exact(a);
    }
    __syncthreads();
    // layer 2
    for (int q = tid; q < kSB * kMlpH; q += kThreads) {
        const int s = q / kMlpH, h = q % kMlpH;
        float a = W[kMlpB2 + h];
        for (int k = 0; k < kMlpH; ++k) a = fmaf(W[kMlpW2 + h * kMlpH + k], h1[s][k], a);
        h2[s][h] = tanh_exact(a);
    }
    __syncthreads();
    // output and loss
    if (tid < kSB) {
        const float t = dy[tid];
        float g = 0.f;
        if (t == t) {
            float y = W[kMlpB3];
            for (int h = 0; h < kMlpH; ++h) y = fmaf(W[kMlpW3 + h], h2[tid][h], y);
            const float r = y - t;
            atomicAdd(A.loss, 0.5f * r * r / (float)A.batch);
            g = r / (float)A.batch;
        }
        dy[tid] = g;
    }
    __syncthreads();
    for (int q = tid; q < kSB * kMlpH; q += kThreads) {
        const int s = q / kMlpH, h = q % kMlpH;
        d2[s][h] = dy[s] * W[kMlpW3 + h] * (1.f - h2[s][h] * h2[s][h]);
    }
    __syncthreads();
    for (int q = tid; q < kSB * kMlpH; q += kThreads) {
        const int s = q / kMlpH, k = q % kMlpH;
        float a = 0.f;
        for (int h = 0; h < kMlpH; ++h) a = fmaf(W[kMlpW2 + h * kMlpH + k], d2[s][h], a);
        d1[s][k] = a * (1.f - h1[s][k] * h1[s][k]);
    }
    __syncthreads();
    // weight gradients summed over the block's samples
    for (int p = tid; p < kMlpNW; p += kThreads) {
        float g = 0.f;
        if (p < kMlpB1) {
            const int h = (p - kMlpW1) / kMlpIn, i = (p - kMlpW1) % kMlpIn;
            for (int s = 0; s < kSB; ++s) g = fmaf(d1[s][h], xn[s][i], g);
        } else if (p < kMlpW2) {
            const int h = p - kMlpB1;
            for (int s = 0; s < kSB; ++s) g += d1[s][h];
        } else if (p < kMlpB2) {
            const int h = (p - kMlpW2) / kMlpH, k = (p - kMlpW2) % kMlpH;
            for (int s = 0; s < kSB; ++s) g = fmaf(d2[s][h], h1[s][k], g);
        } else if (p < kMlpW3) {
            const int h = p - kMlpB2;
            for (int s = 0; s < kSB; ++s) g += d2[s][h];
        } else if (p < kMlpB3) {
            const int h = p - kMlpW3;
            for (int s = 0; s < kSB; ++s) g = fmaf(dy[s], h2[s][h], g);
        } else {
            for (int s = 0; s < kSB; ++s) g += dy[s];
        }
        atomicAdd(&A.grad[p], g);
    }
}

__global__ void mlp_adam_kernel(MlpAdamArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpNW) return;
    const float g = A.grad[p];
    const float m = A.b1 * A.m[p] + (1.f - A.b1) * g;
    const float v = A.b2 * A.v[p] + (1.f - A.b2) * g * g;
    A.m[p] = m;
    A.v[p] = v;
    const float mh = m / (1.f - A.b1t), vh = v / (1.f - A.b2t);
    const float w = A.w[p] - A.lr * mh / (sqrtf(vh) + A.eps);
    A.w[p] = w;
    A.w16[p] = __float2half(w);
    A.grad[p] = 0.f;
}

__global__ void mlp_to_half_kernel(const float *w, __half *w16, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) w16[p] = __float2half(w[p]);
}

// The tracker's inference on given features: one wave per 8 samples, sample g of the wave in
// group g, exactly as the tracker evaluates pending proposals (normalisation, mlp_eval_wave).
__global__ __launch_bounds__(64) void mlp_eval_kernel(StepModel M, const float *x, int64_t n, float *y) {
    __shared__ float area[8][128];
    const int role = threadIdx.x & 7, g = threadIdx.x >> 3;
    const int64_t i = (int64_t)blockIdx.x * 8 + g;
    const bool valid = i < n;
    if (valid) area[g][role] = (x[i * kMlpIn + role] - M.xmu[role]) / M.xsd[role];
    const uint64_t lanes = __ballot(valid && role == 0);
    const MlpLane lw = mlp_lane_weights(M);
    const float out = mlp_eval_wave(M, lw, &area[0][0], 128, lanes, g, role);
    if (valid && role == 0) y[i] = out;
}

hipError_t launch_mlp_eval(const StepModel &m, const float *x, int64_t n, float *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_eval_kernel, dim3((unsigned)((n + 7) / 8)), dim3(64), 0, s, m, x, n, y);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const MlpGradArgs &a, hipStream_t s) {
    const int blocks = (int)((a.batch + kSB - 1) / kSB);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpAdamArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpNW + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_to_half(const float *w, __half *w16, hipStream_t s) {
    hipLaunchKernelGGL(mlp_to_half_kernel, dim3((kMlpNW + 255) / 256), dim3(256), 0, s, w, w16, kMlpNW);
    return hipGetLastError();
}

}  // namespace frg
