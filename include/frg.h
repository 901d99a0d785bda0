/*
 * frg.h -- C ABI of the MI355X homotopy path tracker for Stephenson II finite root generation.
 *
 * Drop-in for the device side of mplecnik/Finite_Root_Generation: the reference's round body
 *     InitRandStates<<<80,256>>>(d_randStates, seed);                               frg.cu:361
 *     Homotopy<<<80,256>>>(d_randStates, d_fpm, d_fpt, d_errorCodes, ...);          frg.cu:374
 *     cudaMemcpy(fpt, d_fpt, ...); cudaMemcpy(errorCodes, d_errorCodes, ...);        frg.cu:389-404
 * (kernels at frg.cu:30-33 and frg.cu:35-124) becomes one call of frg_track().  Inputs and
 * outputs keep the reference's layouts: the target parameters are the 30 complex values of
 * Vect<nP> (frg.cu:149-178), each root is the 21 complex values of Vect<nV> (336 bytes, the
 * record written to the fptColl file, frg.cu:472-476) and each error code is an int32 in 0..5
 * (htpytrack.cu:85-216).  No HIP or torch types cross this boundary.
 *
 * Conventions: functions return FRG_OK (0) or a negative FRG_E* code; per-path outcomes are in
 * the err[] output.  A call takes at most FRG_MAX_COUNT paths, roots or records (FRG_EINVAL
 * beyond); a failed host allocation returns FRG_ENOMEM, and no C++ exception leaves the library.  The caller owns every host buffer; a context owns its device memory and
 * stream.  A context is used by one host thread at a time; separate contexts are independent.
 * Work that an entry point leaves running on a caller's stream (frg_track_device,
 * frg_sort_dedupe_merge_device) uses the context's buffers: the next call on the same context,
 * whatever its stream, first waits for that work to finish (frg_last_kernel_ms excepted).
 */
#ifndef FRG_H
#define FRG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRG_NV 21 /* variables, htpykinematics.cuh:8 */
#define FRG_NP 30 /* parameters, htpykinematics.cuh:9 */
#define FRG_NY 23 /* homogeneous path point [z, Z0, t] */
#define FRG_MAX_COUNT ((int64_t)1 << 36) /* paths, roots or records per call */

enum {
    FRG_OK = 0,
    FRG_EINVAL = -1,      /* bad argument */
    FRG_EHIP = -2,        /* HIP runtime error (device, launch, copy) */
    FRG_ENOMEM = -3,      /* device or host allocation failed */
    FRG_EUNSUPPORTED = -4, /* option combination not implemented */
    FRG_ELIMIT = -5        /* size beyond the fixed limit of an entry point */
};

enum { FRG_START_UNIFORM = 0, FRG_START_GAUSSIAN = 1 };          /* start draws, frgkinematics.cu:7-8 */
enum { FRG_PREC_FP64 = 0, FRG_PREC_MIXED = 1 };                  /* arithmetic of the tracker */
enum { FRG_SCHED_STATIC = 0, FRG_SCHED_PERSISTENT = 1 };          /* one path per group / work queue */
enum { FRG_STEP_CLASSIC = 0, FRG_STEP_LEARNED = 1 };              /* step-size proposal */

/* One complex number, bit-compatible with the reference's Cplx (classes.cuh:26-27). */
typedef struct {
    double re, im;
} frg_c64;

/* Tracker settings.  frg_default_opts() fills the reference's #defines (htpytrack.cu:9-31,
 * frgkinematics.cu:7-8), the FP64 / static / classic choices and a 2-iteration final polish. */
typedef struct {
    double sinit, smax, smin, rk_tol;  /* htpytrack.cu:9-12 */
    int32_t caution_steps, newt_its;   /* :13, :16 */
    double newt_tol;                   /* :17 */
    int32_t newt_attempts, sharp_its;  /* :19, :20 */
    double sharp_tol;                  /* :21 */
    int32_t sharp_attempts, newt_residual; /* :23, :18 (only 1 = |Ydiff| is supported) */
    double sharp_trigger, nudge_trigger;   /* :24, :25 */
    int32_t max_steps, max_ministeps;      /* :26, :27 */
    double patch_lo, patch_hi;             /* :30-31 */
    double start_lo, start_hi;             /* frgkinematics.cu:7-8 */
    int32_t start_mode;                    /* FRG_START_* */
    int32_t precision;                     /* FRG_PREC_* */
    double fp64_band;                      /* mixed: FP64 once 1 - t < fp64_band */
    int32_t schedule;                      /* FRG_SCHED_* */
    int32_t block_size;                    /* threads per block (64..256, multiple of 64); 0 = tuned */
    int32_t groups_per_cu;                 /* persistent grid: paths in flight per CU; 0 = occupancy */
    int32_t step_control;                  /* FRG_STEP_* */
    int32_t use_graph;                     /* 1: launch through a cached hipGraph (queue reset + tracker) */
    int32_t lanes_per_path;                /* lanes of a wave64 that track one path: 8 (8 paths per wave) or 16
                                              (4 paths per wave, FP64 only); 0 = default (8) */
    int32_t polish_its;                    /* FP64 Newton iterations on a sharpened root in the dehomogenising
                                              patch e_21 (frg.cu:44-46) at t = 1; 0 = none, as the reference */
    int32_t waves_per_simd;                /* register budget of the 16-lane tracker as resident waves per SIMD:
                                              1 (up to 512 VGPRs per lane) or 2 (256); 0 = 1.  Only 1 with 8 lanes */
    int32_t pad0;
    double ml_safety;                      /* learned step control: the model's step is exp(prediction) times this
                                              factor (default 0.6) */
    double ml_growth;                      /* learned step control: the model's step is at most this multiple of
                                              the current step (default 5, the classical rule's cap) */
} frg_opts;

/* Per-call summary. */
typedef struct {
    int64_t n_paths;
    int64_t code_count[8];  /* paths per error code 0..5 */
    int64_t steps, rk_attempts, newton_calls, newton_its, tangent_solves;
    double kernel_ms;       /* device time of the tracking launch (HIP events) */
    double total_ms;        /* host wall time of the whole call, copies included */
    int32_t block_size, grid_size, schedule, precision;
    double refine_ms;       /* device time of the refinement + cognate launch that follows it */
    int32_t lanes_per_path, waves_per_simd;
} frg_stats;

/* Opaque context: device, stream, device buffers, tuning cache. */
typedef struct frg_ctx frg_ctx;

/* Fills the defaults.  No reference counterpart (the reference uses #defines). */
int frg_default_opts(frg_opts *opts);

/* Creates a context on HIP device `device`.  Replaces the allocations and setup of
 * frg.cu:138-236 (device properties, cudaMalloc of the round buffers). */
int frg_create(frg_ctx **out, int device, const frg_opts *opts);

/* Frees the context.  Replaces frg.cu:518-531. */
void frg_destroy(frg_ctx *ctx);

/* Replaces the opts of an existing context. */
int frg_set_opts(frg_ctx *ctx, const frg_opts *opts);

/* Current opts of a context (after frg_autotune: the chosen launch shape). */
int frg_get_opts(const frg_ctx *ctx, frg_opts *opts);

/* One round: n_paths paths, path i with RNG id first_path_id + i, random start systems
 * (GenerateStart, frgkinematics.cu:13-136), gamma trick to the target fpm.  Host buffers:
 * fpt (n x 21), err (n), path_stats (n x 8 int32: steps, RK attempts, Newton calls, Newton
 * iterations, ministeps, patch re-draws, tangent solves, solves done in FP32; may be NULL),
 * summary (may be NULL).  Tangent solves are the reference's IVP calls less one per retried
 * rejection: a retry reuses the rejected attempt's first stage tangent, which is bit-identical.  Replaces frg.cu:356-404 (InitRandStates + Homotopy + copies). */
int frg_track(frg_ctx *ctx, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
              int64_t n_paths, frg_c64 *fpt, int32_t *err, int32_t *path_stats, frg_stats *summary);

/* The same with device pointers (no host copies); output buffers must hold n_paths records.
 * `stream` is a hipStream_t (NULL: the context's stream).  The call returns after the work is
 * enqueued; synchronise on the stream before reading.  The next call on the context waits for
 * this launch (it reads the context's instance buffers and work-queue counter). */
int frg_track_device(frg_ctx *ctx, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                     int64_t n_paths, frg_c64 *d_fpt, int32_t *d_err, int32_t *d_path_stats, void *stream);

/* Explicit start systems: spt (n x 21) solves F(spt; spm) = 0 for spm (n x 30).  TrackPath's
 * patch draws for path i come from RNG id first_path_id + i, from draw 0.  yfinal (n x 23, may
 * be NULL) receives the homogeneous end points before dehomogenisation.  Replaces the
 * GenerateStart call in the Homotopy kernel (frg.cu:54) by caller-supplied data. */
int frg_track_explicit(frg_ctx *ctx, const frg_c64 *fpm, frg_c64 gamma, const frg_c64 *spt, const frg_c64 *spm,
                       uint64_t seed, uint64_t first_path_id, int64_t n_paths, frg_c64 *fpt, int32_t *err,
                       int32_t *path_stats, frg_c64 *yfinal, frg_stats *summary);

/* Several task instances in one launch: instance k has target fpm[30k..30k+29], gamma[k],
 * seed seeds[k], RNG ids first_path_ids[k] + i, and paths_per_inst[k] paths; outputs are
 * concatenated in instance order.  No reference counterpart (the reference tracks one target). */
int frg_track_batch(frg_ctx *ctx, int32_t n_inst, const frg_c64 *fpm, const frg_c64 *gamma, const uint64_t *seeds,
                    const uint64_t *first_path_ids, const int64_t *paths_per_inst, frg_c64 *fpt, int32_t *err,
                    int32_t *path_stats, frg_stats *summary);

/* frg_track_batch on device buffers without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  d_fpm (n_inst x 30), d_fpt (total x 21), d_err
 * (total) and d_path_stats (total x 8, may be NULL) are device pointers; gamma, seeds,
 * first_path_ids and paths_per_inst are host arrays of n_inst entries, passed to the GPU as kernel
 * arguments.  The work-queue counter and the instance descriptions live in a fixed scratch array
 * of FRG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: n_inst is at most
 * FRG_NOALLOC_MAX_INSTANCES (FRG_ELIMIT beyond).  Launches eagerly on `stream` (NULL: the
 * context's stream) whatever opts.use_graph says and returns without waiting.  Every call on one
 * GPU, from any context, shares that array and the call overwrites it: calls that overlap on one
 * GPU race.  Results are those of frg_track_batch with the same opts. */
#define FRG_NOALLOC_MAX_INSTANCES 64
#define FRG_NOALLOC_SCRATCH_BYTES 4096
int frg_track_batch_noalloc(frg_ctx *ctx, int32_t n_inst, const frg_c64 *d_fpm, const frg_c64 *gamma,
                            const uint64_t *seeds, const uint64_t *first_path_ids, const int64_t *paths_per_inst,
                            frg_c64 *d_fpt, int32_t *d_err, int32_t *d_path_stats, void *stream);

/* One measured launch shape of frg_autotune. */
typedef struct {
    int32_t schedule, block_size, groups_per_cu; /* groups_per_cu: resident paths per CU (persistent) */
    int32_t lanes_per_path;                      /* lanes per path: paths per wave = 64 / lanes_per_path */
    int32_t waves_per_simd, pad0;                /* register budget (frg_opts.waves_per_simd) */
    double kernel_ms, paths_per_s;
} frg_tune_entry;

/* Times the tracking kernel on n_sample paths (RNG ids first_path_id + i of `seed`, target
 * fpm/gamma) for each candidate launch shape -- paths per wave (8 lanes per path, or 16 for FP64
 * with a register budget of 1 or 2 waves per SIMD),
 * static or persistent schedule, 64/128/256-thread blocks, and for the persistent grid 1..max
 * blocks per CU -- with the context's precision and step control, then sets the context's
 * lanes_per_path, waves_per_simd, schedule, block_size and groups_per_cu to the fastest.  Up to table_cap
 * measurements go to table (may be NULL), their number to *n_entries.  A sample of a few hundred
 * thousand paths keeps the launch's tail (its slowest paths) a small part of each timing.  No
 * reference counterpart (the reference launches 80 x 256, frg.cu:25-27). */
int frg_autotune(frg_ctx *ctx, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                 int64_t n_sample, frg_tune_entry *table, int32_t table_cap, int32_t *n_entries);

/* Start systems alone (GenerateStart, frgkinematics.cu:13-136): spt (n x 21), spm (n x 30). */
int frg_generate_starts(frg_ctx *ctx, uint64_t seed, uint64_t first_path_id, int64_t n, frg_c64 *spt, frg_c64 *spm);

/* One linear solve per path on explicit inputs, for stage-level parity:
 *   mode 0: IVP with tstepQ = false (htpysupport.cu:237-299), w = velocity patch row;
 *   mode 1: IVP with tstepQ = true;
 *   mode 2: one Newton update toward tgoal[i] (htpysupport.cu:430-480; w ignored).
 * Y, u, w: n x 23; spm: n x 30; fpm: 30; x: n x 23 out.  precision FRG_PREC_FP64 or
 * FRG_PREC_MIXED (FP32 Jacobian and solve, FP64 residual).  The solve is the tracker's for the
 * context's lanes_per_path (16: FP64 only, FRG_EUNSUPPORTED otherwise). */
int frg_solve_step(frg_ctx *ctx, int32_t mode, int32_t precision, int64_t n, const frg_c64 *Y, const frg_c64 *u,
                   const frg_c64 *w, const frg_c64 *spm, const frg_c64 *fpm, frg_c64 gamma, const double *tgoal,
                   frg_c64 *x);

/* The step-size factor of an RK attempt as the tracker computes it, for explicit error
 * estimates err[0..n-1]: scale = 0.8 (rk_tol/err)^(1/6), at most 5, and 5 for err <= 1e-16
 * (rungekutta.cu:107-112) with the context's rk_tol; stage-level parity entry point. */
int frg_step_factor(frg_ctx *ctx, int64_t n, const double *err, double *scale);

/* Host post-processing (frgsupport.cu).  frg_sort_dedupe_merge replaces SortingDeletingMerging
 * (frgsupport.cu:10-215) plus the device sort of frg.cu:432-435: merges the n new roots into
 * the sorted collection coll (m roots), marks duplicates in err with code 6 (tolerance tol on
 * every real, frgsupport.cu:7), writes the merged sorted collection to coll_out (room for
 * m + n) and returns its length (negative on error).  frg_lambertw replaces LambertW
 * (frgsupport.cu:249-272); frg_completeness gives the estimated fraction of roots collected,
 * 1 + r W0(-e^{-1/r}/r) with r = successes/(successes + duplicates) (frg.cu:464-465). */
int64_t frg_sort_dedupe_merge(const frg_c64 *coll, int64_t m, const frg_c64 *fpt, int32_t *err, int64_t n,
                              double tol, frg_c64 *coll_out);
/* frg_sort_dedupe_merge on the device, with the same rule and a bit-identical result: all four
 * arrays are device pointers (d_coll_out: room for m + n roots), m + n < 2^31, *m_out receives
 * the merged length.  Replaces the device sort of frg.cu:432-435 (thrust::sort_by_key) and the
 * host SortingDeletingMerging (frgsupport.cu:10-215).  `stream` is a hipStream_t (NULL: the
 * context's stream); the call synchronises it before returning. */
int frg_sort_dedupe_merge_device(frg_ctx *ctx, const frg_c64 *d_coll, int64_t m, const frg_c64 *d_fpt,
                                 int32_t *d_err, int64_t n, double tol, frg_c64 *d_coll_out, int64_t *m_out,
                                 void *stream);
double frg_lambertw(double z);
double frg_completeness(int64_t successes, int64_t duplicates);

/* ---- Learned step-size control (no reference counterpart: the reference's controller is the
 * classical rule of rungekutta.cu:109-112 and htpytrack.cu:179-182, which stays the fallback).
 * A model is an 8 -> 64 -> 64 -> 1 tanh MLP (4801 weights) plus input/output normalisation:
 *   W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3, xmu[8], xsd[8], ymu, ysd  (FP32)
 * A step record is 8 features of a step proposal and the log of the step that met the RK
 * tolerance on the next accepted step. */
#define FRG_STEP_MODEL_SIZE 4819
#define FRG_STEP_MODEL_WEIGHTS 4801
#define FRG_STEP_RECORD 9

/* Tracks n_paths paths (as frg_track, outputs discarded) and returns the step records whose
 * outcome was observed, at most cap (cap <= INT32_MAX); *n_records receives their number. */
int frg_collect_step_log(frg_ctx *ctx, const frg_c64 *fpm, frg_c64 gamma, uint64_t seed, uint64_t first_path_id,
                         int64_t n_paths, float *records, int64_t cap, int64_t *n_records);

/* Trains a model on n records with the GPU kernels (minibatch Adam, `steps` steps of `batch`
 * records drawn with replacement from seed).  model: FRG_STEP_MODEL_SIZE floats out; loss
 * (may be NULL): the minibatch loss of every step. */
int frg_train_step_model(frg_ctx *ctx, const float *records, int64_t n, int32_t steps, int32_t batch, float lr,
                         uint64_t seed, float *model, float *loss);

/* Loads a model for FRG_STEP_LEARNED tracking (weights stored as FP16 on the device). */
int frg_set_step_model(frg_ctx *ctx, const float *model);

/* The tracker's own inference (FP16 weights, FP32 accumulation, the kernel code that proposes
 * steps) on n feature vectors (n x 8 floats, unnormalised, as in a step record) with the loaded
 * model; out receives the n predicted log step sizes.  No reference counterpart: a parity entry
 * point for the inlined inference. */
int frg_step_model_eval(frg_ctx *ctx, const float *features, int64_t n, float *out);

/* Gradient of the mean loss over records 0..n-1 with respect to the 4801 weights (GPU). */
int frg_step_model_grad(frg_ctx *ctx, const float *model, const float *records, int64_t n, float *grad, float *loss);

/* Device time (ms, HIP events) of the last tracking launch of the context; waits for it to end.
 * A launch is the tracking kernel followed by the refinement kernel (FP64 polish of code-0 roots
 * with an extended residual, then CognateSolns, frgkinematics.cu:227-285); frg_last_kernel_ms
 * times the first (for a graph launch: both), frg_last_refine_ms the second.  No reference
 * counterpart (the reference times RNG init + kernel + copies, frg.cu:356-426). */
int frg_last_kernel_ms(frg_ctx *ctx, double *ms);
int frg_last_refine_ms(frg_ctx *ctx, double *ms);

/* Algorithmic floating-point operations per unit of work of this implementation's formulation
 * (real +, -, *, / count 1; counted once per path, not per lane).  A tracked path costs
 *   tangent_solves * tangent_solve + newton_its * newton_update + rk_attempts * rk_combination
 * (counters from path_stats).  No reference counterpart. */
typedef struct {
    double pose_rows;       /* one pose: values, Jacobian rows and dF/dt of its three equations */
    double pose_values;     /* one pose: the three residuals only */
    double tangent_solve;   /* one tangent (IVP) solve, normalisation included */
    double newton_update;   /* one Newton iteration */
    double rk_combination;  /* stage points, error estimate and solution of one RK attempt */
} frg_flop_model;
int frg_flops(frg_flop_model *out);

/* Error text for a return code. */
const char *frg_strerror(int code);

/* ABI version (major * 10000 + minor * 100 + patch). */
int frg_version(void);

#ifdef __cplusplus
}
#endif
#endif
