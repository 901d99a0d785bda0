/*
 * frg_oracle.h -- CPU oracle for the Stephenson II finite-root-generation path tracker.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's per-path
 * routines (mplecnik/Finite_Root_Generation: classes.cuh, support.cu, htpykinematics.cu,
 * htpysupport.cu, rungekutta.cu, htpytrack.cu, frgkinematics.cu, the Homotopy kernel body
 * frg.cu:35-69, and the host post-processing in frgsupport.cu).  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it, and only to compare against: the product path
 * (finite_root_generation_amd/) never links or calls it.
 *
 * Random numbers: the reference draws from curand XORWOW, which is not reproduced.  Both this
 * oracle and the GPU tracker draw from Philox4x32-10 (Salmon et al., SC'11) keyed by
 * (seed, path id); draw k of a path is half (k & 1) of Philox block k >> 1.  The reference's
 * quirk that GenerateStart and TrackPath both start from a copy of the same generator state
 * (frg.cu:54,57) is kept: both read the path's stream from draw 0.
 *
 * Layout: every complex number is {double re, im} (bit-compatible with the reference's Cplx);
 * Y vectors have 23 entries [z(21), Z0, t]; parameter vectors have 30 entries
 * [P0..P7, Pc0..Pc7, Q1..Q7, Qc1..Qc7] (htpykinematics.cu:29-33).
 */
#ifndef FRG_ORACLE_H
#define FRG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ONV 21 /* number of variables, htpykinematics.cuh:8 */
#define ONP 30 /* number of parameters, htpykinematics.cuh:9 */
#define ONY 23 /* [z, Z0, t] */

typedef struct { double re, im; } ocplx;

/* Tracker settings; the defaults (oracle_default_opts) are the #defines of htpytrack.cu:9-31
 * and frgkinematics.cu:7-8. */
typedef struct {
    double sinit, smax, smin, rk_tol;
    int32_t caution_steps, newt_its;
    double newt_tol;
    int32_t newt_attempts, sharp_its;
    double sharp_tol;
    int32_t sharp_attempts, newt_residual;
    double sharp_trigger, nudge_trigger;
    int32_t max_steps, max_ministeps;
    double patch_lo, patch_hi;
    double start_lo, start_hi;
    int32_t start_mode;     /* 0: U[start_lo,start_hi]^2 (reference), 1: standard complex Gaussian */
    int32_t sharp_residual; /* 1: |Ydiff| (reference), 0: |H| */
    int32_t polish_its;     /* Newton iterations on a code-0 root in the dehomogenising patch e_21
                               (frg.cu:44-46) at t = 1 after sharpening; 0 = none, as the
                               reference.  Not in the reference: the final FP64 refinement. */
    int32_t pad0;
} otrk_opts;

/* Per-path counters (not in the reference; used for FLOP accounting and histograms). */
typedef struct {
    int32_t steps;         /* accepted steps (the reference's `step`) */
    int32_t rk_attempts;   /* DormandPrince65 calls */
    int32_t newton_calls;  /* Newton calls (corrector + sharpening) */
    int32_t newton_its;    /* Newton iterations, all calls */
    int32_t ministeps;     /* the reference's `ministep` */
    int32_t redraws;       /* patch re-draws */
    int32_t ivp_calls;     /* IVP calls (8 per RK attempt + 1) */
    int32_t end_code;      /* error code, duplicated for convenience */
    int32_t rejections;    /* rejected RK attempts followed by a retry */
} otrk_stats;

void oracle_default_opts(otrk_opts *o);

/* ---- RNG ---- */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double oracle_uniform(uint64_t seed, uint64_t path_id, uint64_t draw);

/* ---- L0 complex primitives (classes.cuh:26-45) ---- */
void oracle_cdiv(const ocplx *a, const ocplx *b, ocplx *out);
void oracle_csqrt(const ocplx *x, ocplx *out);

/* ---- L1 system (htpykinematics.cu) ---- */
void oracle_fsys(const ocplx *Z /*22*/, const ocplx *q /*30*/, ocplx *F /*21*/);
/* FSYS evaluated in IEEE binary128 and rounded to double (the refinement's residual). */
void oracle_fsys_quad(const ocplx *Z /*22*/, const ocplx *q /*30*/, ocplx *F /*21*/);
void oracle_jvar(const ocplx *Z, const ocplx *q, ocplx *J /*21x22 row-major*/);
void oracle_jpar(const ocplx *Z, const ocplx *q, ocplx *J /*21x30 row-major*/);

/* ---- L2 homotopy support (htpysupport.cu) ---- */
void oracle_switch_patch(const ocplx *unew, const ocplx *Yold, ocplx *Ynew);
void oracle_q(double t, const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *out);
void oracle_dqdt(double t, const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *out);
void oracle_H(const ocplx *Y, const ocplx *u, const ocplx *spm, const ocplx *fpm,
              const ocplx *gamma, ocplx *out /*22*/);
void oracle_JH(const ocplx *Y, const ocplx *u, const ocplx *spm, const ocplx *fpm,
               const ocplx *gamma, ocplx *out /*22x23*/);
void oracle_linear_solve(ocplx *mat /*23x23, destroyed*/, ocplx *b /*23, destroyed*/, ocplx *x);
void oracle_linear_solve_newton(ocplx *mat, ocplx *b /*22*/, double tdiff, ocplx *x /*23*/);
void oracle_ivp(int tstep, const ocplx *Yc, const ocplx *Vc, const ocplx *u, const ocplx *spm,
                const ocplx *fpm, const ocplx *gamma, ocplx *dYds);
int oracle_newton(double tgoal, const ocplx *Vn, const ocplx *u, const ocplx *spm, const ocplx *fpm,
                  const ocplx *gamma, int max_its, double tol, int residual_type, ocplx *YN,
                  double *residual);

/* ---- L3 integrator (rungekutta.cu) ---- */
void oracle_rk65(double s, double tol, int tstep, const ocplx *Yc, const ocplx *Vc, const ocplx *u,
                 const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *Yn, ocplx *Vn,
                 double *err, double *scale);

/* ---- L4 tracker (htpytrack.cu) ---- */
int oracle_track_path(const otrk_opts *o, const ocplx *spt, const ocplx *spm, const ocplx *fpm,
                      const ocplx *gamma, uint64_t seed, uint64_t path_id, uint64_t first_draw,
                      ocplx *Yfinal /*23*/, otrk_stats *st);

/* The final refinement alone, on a dehomogenised point Y (23: z, Z0 := 1, t = 1): polish_its
 * Newton steps in the patch e_21 with the binary128 residual.  Returns the steps taken. */
int oracle_refine_root(const otrk_opts *o, const ocplx *fpm /*30*/, ocplx gamma, ocplx *Y /*23*/);

/* ---- L5 start system and cognates (frgkinematics.cu) ---- */
void oracle_generate_start(const otrk_opts *o, uint64_t seed, uint64_t path_id, ocplx *z /*21*/,
                           ocplx *q /*30*/);
void oracle_cognate12(const ocplx *orig /*7*/, ocplx P, ocplx *cogn /*7*/);
void oracle_cognate_solns(const ocplx *z /*21*/, const ocplx *q /*30*/, ocplx *out /*2x21*/);

/* ---- L6 kernel body (frg.cu:35-69), many paths with OpenMP ---- */
/* Generated starts: path i uses id first_path_id+i.  yfinal (n x 23, optional) receives the
 * homogeneous end point before dehomogenisation. */
int oracle_homotopy(const otrk_opts *o, const ocplx *fpm, ocplx gamma, uint64_t seed,
                    uint64_t first_path_id, int32_t n, ocplx *fpt, int32_t *err, otrk_stats *stats,
                    ocplx *yfinal, int nthreads);
/* Explicit starts spt (n x 21) and spm (n x 30); TrackPath's draws start at draw 0. */
int oracle_homotopy_explicit(const otrk_opts *o, const ocplx *fpm, ocplx gamma, const ocplx *spt,
                             const ocplx *spm, uint64_t seed, uint64_t first_path_id, int32_t n,
                             ocplx *fpt, int32_t *err, otrk_stats *stats, ocplx *yfinal,
                             int nthreads);

/* Explicit starts with an RNG path id per path (patch draws of path i from path_ids[i]): many
 * copies of one path under perturbed start data in one parallel call. */
int oracle_homotopy_explicit_ids(const otrk_opts *o, const ocplx *fpm, ocplx gamma, const ocplx *spt,
                                 const ocplx *spm, uint64_t seed, const uint64_t *path_ids, int32_t n,
                                 ocplx *fpt, int32_t *err, int nthreads);

/* ---- L7 host post-processing (frgsupport.cu) ---- */
double oracle_lambertw(double z);
/* Merge n new roots into a sorted collection of m roots (both AoS 21 complex). Duplicates get
 * code 6 in err.  coll_out must hold m + n roots; returns the merged count. */
int64_t oracle_sort_dedupe_merge(const ocplx *coll, int64_t m, const ocplx *fpt, int32_t *err,
                                 int64_t n, double tol, ocplx *coll_out);

#ifdef __cplusplus
}
#endif
#endif
