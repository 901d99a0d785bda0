// frg_kernels.h -- kernel arguments shared by the tracker kernels and the host C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frg_device.h"
#include "frg_mlp.h"

namespace frg {

// Tracker settings (defaults: htpytrack.cu:9-31, frgkinematics.cu:7-8).
struct TrackOpts {
    double sinit, smax, smin, rk_tol;
    int caution_steps, newt_its;
    double newt_tol;
    int newt_attempts, sharp_its;
    double sharp_tol;
    int sharp_attempts;
    int polish_its;     // FP64 Newton iterations in the dehomogenising patch after sharpening
    double sharp_trigger, nudge_trigger;
    int max_steps, max_ministeps;
    double patch_lo, patch_hi, start_lo, start_hi;
    int start_mode;     // 0: U[start_lo, start_hi]^2 (reference), 1: standard complex Gaussian
    int precision;      // 0: FP64, 1: FP32 Jacobian/solve away from t = 1
    double fp64_band;   // mixed precision: FP64 once 1 - t < fp64_band
    double ml_safety;   // learned step control: factor on the model's step
    double ml_growth;   // learned step control: cap on the model's step as a multiple of the current one
};

// One launch.  Several instances (task targets) may share a launch; instance i owns global path
// indices [inst_first[i], inst_first[i+1]).
struct Job {
    int64_t n_total;
    int n_inst;
    const int64_t *inst_first;    // n_inst + 1 (may be null when n_inst == 1)
    const cx<double> *fpm;        // n_inst x 30 target parameters
    const cx<double> *gamma;      // n_inst
    const uint64_t *seed;         // n_inst RNG seeds
    const uint64_t *pid0;         // n_inst: RNG path id of the instance's first path
    const cx<double> *spt, *spm;  // explicit start systems (n_total x 21, n_total x 30) or null
    cx<double> *fpt;              // n_total x 21 output roots (AoS, the reference's Vect<21>)
    int *err;                     // n_total error codes 0..5
    int *stats;                   // n_total x 8 counters or null
    cx<double> *yfinal;           // n_total x 23 homogeneous end points or null
    unsigned long long *queue;    // work-queue head (persistent launches)
    TrackOpts o;
    // learned step control
    int learned;                  // 1: the model proposes the step after an accepted step
    StepModel model;
    // step log: (features of a proposal, log of the step that met the tolerance next)
    float *steplog;               // steplog_cap x kStepRecord, or null
    unsigned long long *steplog_count;
    int64_t steplog_cap;
};

// Stage-level parity entry points (one group solve per path; start generation alone).
struct DebugSolveArgs {
    int64_t n;
    int mode;  // 0 arclength tangent, 1 t-step tangent, 2 Newton update
    const cx<double> *Y, *u, *w, *spm, *fpm;  // n x 23, n x 23, n x 23, n x 30, 30
    const double *tgoal;                       // n (Newton)
    cx<double> gamma;
    cx<double> *x;                             // n x 23 out
};
struct DebugStartArgs {
    int64_t n;
    uint64_t seed, pid0;
    TrackOpts o;
    cx<double> *spt, *spm;  // n x 21, n x 30 out
};
hipError_t launch_debug_solve(const DebugSolveArgs &a, int precision, int lanes, hipStream_t stream);
hipError_t launch_debug_start(const DebugStartArgs &a, hipStream_t stream);
// the tracker's step-size factor for explicit error estimates (n values)
hipError_t launch_debug_step_factor(const double *err, int64_t n, double tol, double *out, hipStream_t stream);

// host-side launchers (frg_tracker.hip)
enum LaunchKind : int { LK_STATIC = 0, LK_PERSISTENT = 1 };
// lanes: lanes of a wave64 per path, 8 (FP64 or mixed) or 16 (FP64); waves: register budget of the
// 16-lane kernel as resident waves per SIMD, 1 (512 VGPRs) or 2 (256); 1 for 8 lanes
hipError_t launch_track(const Job &job, int kind, int block, int grid, int lanes, int waves, hipStream_t stream);
// refinement of code-0 roots and cognate canonicalisation of every output row (after the tracker)
hipError_t launch_refine(const Job &job, hipStream_t stream);
// zeroes a work-queue counter with a kernel (capturable, unlike a memset node)
hipError_t launch_reset_counter(unsigned long long *p, hipStream_t stream);
size_t track_lds_bytes(int block, int lanes);

// Allocation-free batch launches: a fixed per-GPU scratch array (frg_tracker.hip) holds the
// work-queue counter and the descriptions of up to kNoAllocMaxInst instances, staged by a kernel
// whose arguments carry them (no host copy, so the sequence can be captured in a graph).
constexpr int kNoAllocMaxInst = 64;
constexpr size_t kNoAllocScratchBytes = 4096;
struct NoAllocStage {
    int n_inst;
    int64_t inst_first[kNoAllocMaxInst + 1];
    cx<double> gamma[kNoAllocMaxInst];
    uint64_t seed[kNoAllocMaxInst], pid0[kNoAllocMaxInst];
};
struct NoAllocLayout {
    unsigned long long *queue;
    int64_t *inst_first;
    cx<double> *gamma;
    uint64_t *seed, *pid0;
};
// offsets in the scratch array: queue 0, inst_first 256, gamma 1024, seed 2048, pid0 2560
__host__ __device__ inline NoAllocLayout noalloc_layout(unsigned char *base) {
    NoAllocLayout L;
    L.queue = reinterpret_cast<unsigned long long *>(base);
    L.inst_first = reinterpret_cast<int64_t *>(base + 256);
    L.gamma = reinterpret_cast<cx<double> *>(base + 1024);
    L.seed = reinterpret_cast<uint64_t *>(base + 2048);
    L.pid0 = reinterpret_cast<uint64_t *>(base + 2560);
    return L;
}
static_assert(256 + 8 * (kNoAllocMaxInst + 1) <= 1024 && 1024 + 16 * kNoAllocMaxInst <= 2048 &&
                  2048 + 8 * kNoAllocMaxInst <= 2560 && 2560 + 8 * kNoAllocMaxInst <= (int)kNoAllocScratchBytes,
              "scratch layout");
static_assert(sizeof(NoAllocStage) <= 4000, "kernel argument size");
// address of this GPU's scratch array (call with the device current)
hipError_t noalloc_scratch_address(void **ptr);
hipError_t launch_stage_instances(const NoAllocStage &a, void *base, hipStream_t stream);
int track_max_groups_per_cu(int block, int kind, int precision, int lanes, int waves);

}  // namespace frg
