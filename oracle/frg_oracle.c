/*
 * frg_oracle.c -- CPU oracle (test infrastructure only; see frg_oracle.h).
 *
 * Every function below restates one routine of mplecnik/Finite_Root_Generation and cites the
 * file:line it follows.  Where the restatement differs in form from the reference the comment
 * says so and why the result is the same.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off,
 * so no fused multiply-adds are formed and the results do not depend on the host CPU).
 *
 * Pinned by tests/test_oracle.py against: Philox known-answer vectors, numpy complex
 * arithmetic, F(start) = 0 for GenerateStart output, a transcription of the reference's expanded
 * FSYS, finite differences (JVAR, JPAR, dqdt), numpy.linalg.solve (LU solves), the RK6(5)
 * tableau's order conditions, Newton's quadratic convergence, the cognate invariant, scipy's
 * Lambert W and brute-force duplicate detection.
 *
 * Mutation testing: built with -DORACLE_MUTANT=k (tests/test_oracle_mutants.py does, into a
 * temporary directory), the library plants the k-th of the deliberate mistakes marked MUT(k)
 * below -- a swapped operand, a dropped term, a wrong sign or index -- so that the tests can show
 * that one of the checks above catches each.  The Makefile builds no mutant.
 */
#include "frg_oracle.h"

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUT(k) (ORACLE_MUTANT == (k))

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* L0: complex scalar (classes.cuh:26-45, support.cu:5-24)                                    */
/* ------------------------------------------------------------------------------------------ */

static inline ocplx C(double re, double im) { ocplx z = {re, im}; return z; }
static inline ocplx cadd(ocplx a, ocplx b) { return C(a.re + b.re, a.im + b.im); }
static inline ocplx csub(ocplx a, ocplx b) { return C(a.re - b.re, a.im - b.im); }
static inline ocplx cneg(ocplx a) { return C(-a.re, -a.im); }
static inline ocplx cmul(ocplx a, ocplx b) {
    return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* Division by Smith's method (R. L. Smith, "Algorithm 116: Complex division", CACM 5(8), 1962). */
static inline ocplx cdiv(ocplx a, ocplx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return C((a.re + a.im * r) / d, (MUT(17) ? a.im + a.re * r : a.im - a.re * r) / d);
    } else {
        double r = b.re / b.im, d = b.im + b.re * r;
        return C((a.re * r + a.im) / d, (a.im * r - a.re) / d);
    }
}
static inline double cabs_(ocplx a) { return hypot(a.re, a.im); }
/* cplxFma(x, y, d) = x*y + d (classes.cuh:31) */
static inline ocplx cfma(ocplx x, ocplx y, ocplx d) { return cadd(cmul(x, y), d); }
/* Vect::scale(Real) multiplies by {s, 0} (classes.cuh:91-96) */
static inline ocplx crs(double s, ocplx a) { return cmul(C(s, 0.0), a); }
/* cplxPow(x, n) by repeated multiplication (support.cu:5-15) */
static inline ocplx cpow2(ocplx x) { return cmul(x, x); }
static inline ocplx cpow3(ocplx x) { return cmul(cmul(x, x), x); }
/* Principal square root by the half-angle formulas of classes.cuh:39-45 (sign copied from the
 * imaginary part, including a signed zero). */
static inline ocplx csqrt_ref(ocplx x) {
    double r = cabs_(x);
    double cosA = x.re / r;
    ocplx out = C(sqrt(r * (cosA + 1.0) / 2.0), sqrt(r * (1.0 - cosA) / 2.0));
    if (MUT(18) ? !signbit(x.im) : signbit(x.im)) out.im *= -1.0;
    return out;
}

void oracle_cdiv(const ocplx *a, const ocplx *b, ocplx *out) { *out = cdiv(*a, *b); }
void oracle_csqrt(const ocplx *x, ocplx *out) { *out = csqrt_ref(*x); }

/* Vect<M>::norm (classes.cuh:82-88) */
static double vnorm(const ocplx *v, int m) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += v[i].re * v[i].re + v[i].im * v[i].im;
    return sqrt(s);
}
/* Vect<M>::dot, sum from zero in index order (classes.cuh:120-126) */
static ocplx vdot(const ocplx *a, const ocplx *b, int m) {
    ocplx p = C(0, 0);
    for (int i = 0; i < m; ++i) p = cadd(p, cmul(a[i], b[i]));
    return p;
}

/* ------------------------------------------------------------------------------------------ */
/* RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                                      */
/* ------------------------------------------------------------------------------------------ */

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (MUT(1) ? c2 : c0);
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (MUT(1) ? c0 : c2);
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform double in (0, 1]: 53 high bits of a 64-bit word, plus one, times 2^-53. */
double oracle_uniform(uint64_t seed, uint64_t path_id, uint64_t draw) {
    uint64_t blk = draw >> 1;
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)path_id,
                       (uint32_t)(path_id >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    oracle_philox4x32_10(ctr, key, o);
    uint64_t w = (draw & 1) ? (((uint64_t)o[3] << 32) | o[2]) : (((uint64_t)o[1] << 32) | o[0]);
    if (MUT(2)) w = (w << 32) | (w >> 32);
    return (double)((w >> 11) + 1) * 0x1.0p-53;
}

typedef struct { uint64_t seed, pid, next; } orng;
static inline double rng_next(orng *g) { return oracle_uniform(g->seed, g->pid, g->next++); }

void oracle_default_opts(otrk_opts *o) {
    memset(o, 0, sizeof(*o));
    o->sinit = 1.0;              /* htpytrack.cu:9  */
    o->smax = 100.0;             /* htpytrack.cu:10 */
    o->smin = 0.000001;          /* htpytrack.cu:11 */
    o->rk_tol = 0.001;           /* htpytrack.cu:12 */
    o->caution_steps = 2;        /* htpytrack.cu:13 */
    o->newt_its = 10;            /* htpytrack.cu:16 */
    o->newt_tol = 0.00000001;    /* htpytrack.cu:17 */
    o->newt_residual = 1;        /* htpytrack.cu:18 */
    o->newt_attempts = 3;        /* htpytrack.cu:19 */
    o->sharp_its = 30;           /* htpytrack.cu:20 */
    o->sharp_tol = 0.000000000001; /* htpytrack.cu:21 */
    o->sharp_residual = 1;       /* htpytrack.cu:22 */
    o->sharp_attempts = 3;       /* htpytrack.cu:23 */
    o->sharp_trigger = 0.00001;  /* htpytrack.cu:24 */
    o->nudge_trigger = 0.00001;  /* htpytrack.cu:25 */
    o->max_steps = 2000;         /* htpytrack.cu:26 */
    o->max_ministeps = 3000;     /* htpytrack.cu:27 */
    o->patch_lo = -2.0;          /* htpytrack.cu:31 */
    o->patch_hi = 2.0;           /* htpytrack.cu:30 */
    o->start_lo = -5.0;          /* frgkinematics.cu:8 */
    o->start_hi = 5.0;           /* frgkinematics.cu:7 */
    o->start_mode = 0;
    o->polish_its = 2;
}

/* ------------------------------------------------------------------------------------------ */
/* L1: the Stephenson II system (htpykinematics.cu)                                            */
/* ------------------------------------------------------------------------------------------ */

/* Variable slots (htpykinematics.cu:10-27) */
enum { iA = 0, iB, iC, iD, iF, iG, iH, iAc, iBc, iCc, iDc, iFc, iGc, iHc };

/* FSYS, restated term by term from the expanded forms at htpykinematics.cu:36-122:
 *   rows 0..6   I_j   (:38-55)
 *   rows 7..13  II_j  (:57-66)
 *   rows 14..20 III_j (:68-122)
 * The terms are summed in the reference's order. */
#define FS_NAME oracle_fsys
#define FS_IN ocplx
#define FS_OUT ocplx
#define FS_T ocplx
#define FS_LD(x) (x)
#define FS_ST(x) (x)
#define FS_ADD cadd
#define FS_SUB csub
#define FS_NEG cneg
#define FS_MUL cmul
#define FS_DIV cdiv
#define FS_P2 cpow2
#define FS_P3 cpow3
#include "fsys_expanded.inc"
#undef FS_NAME
#undef FS_IN
#undef FS_OUT
#undef FS_T
#undef FS_LD
#undef FS_ST
#undef FS_ADD
#undef FS_SUB
#undef FS_NEG
#undef FS_MUL
#undef FS_DIV
#undef FS_P2
#undef FS_P3

/* The same residuals in IEEE binary128 (gcc's __float128 complex, software arithmetic): the
 * extended-precision residual of the final refinement.  Inputs are doubles, taken exactly. */
typedef _Complex float __attribute__((mode(TC))) qcplx;
static inline qcplx q_of(ocplx z) {
    qcplx r;
    __real__ r = (__float128)z.re;
    __imag__ r = (__float128)z.im;
    return r;
}
static inline ocplx o_of(qcplx z) { return C((double)__real__ z, (double)__imag__ z); }
static inline qcplx qadd(qcplx a, qcplx b) { return a + b; }
static inline qcplx qsub(qcplx a, qcplx b) { return a - b; }
static inline qcplx qneg(qcplx a) { return -a; }
static inline qcplx qmul(qcplx a, qcplx b) { return a * b; }
static inline qcplx qdiv(qcplx a, qcplx b) { return a / b; }
static inline qcplx qp2(qcplx a) { return a * a; }
static inline qcplx qp3(qcplx a) { return a * a * a; }
#define FS_NAME fsys_quad
#define FS_IN ocplx
#define FS_OUT ocplx
#define FS_T qcplx
#define FS_LD q_of
#define FS_ST o_of
#define FS_ADD qadd
#define FS_SUB qsub
#define FS_NEG qneg
#define FS_MUL qmul
#define FS_DIV qdiv
#define FS_P2 qp2
#define FS_P3 qp3
static
#include "fsys_expanded.inc"
#undef FS_NAME
#undef FS_IN
#undef FS_OUT
#undef FS_T
#undef FS_LD
#undef FS_ST
#undef FS_ADD
#undef FS_SUB
#undef FS_NEG
#undef FS_MUL
#undef FS_DIV
#undef FS_P2
#undef FS_P3

void oracle_fsys_quad(const ocplx *Z, const ocplx *q, ocplx *F) { fsys_quad(Z, q, F); }

/* All three equation families share one shape (derived from the expanded forms above):
 *     F = Z0*K - W*Wc,   W = Z0*X - U*Y,   Wc = Xc - (Z0/U)*Yc,
 * with X, Y (resp. Xc, Yc) affine in the unconjugated (resp. conjugated) linkage variables and
 * K free of U and Z0.  JVAR and JPAR (htpykinematics.cu:129-1384) are the exact partial
 * derivatives of FSYS; here they are computed from this shape by the product rule:
 *   d/dv  (v unconjugated) = Z0*dK - (Z0*dX - U*dY)*Wc
 *   d/dvc (v conjugated)   = Z0*dK - W*(dXc - (Z0/U)*dYc)
 *   d/dU  = Y*Wc - W*(Z0/U)*Yc/U
 *   d/dZ0 = K - (X + Z0*dX/dZ0 - U*dY/dZ0)*Wc - W*(dXc/dZ0 - Yc/U - (Z0/U)*dYc/dZ0).
 * The formula differs from the reference's machine-generated expression trees; the values
 * agree to rounding (tests compare against finite differences of the expanded FSYS). */
typedef struct {
    ocplx Z0, U, zu, W, Wc, X, Y, Xc, Yc, K;
} eqshape;

static inline void shape_finish(eqshape *e) {
    e->W = csub(cmul(e->Z0, e->X), cmul(e->U, e->Y));
    e->Wc = csub(e->Xc, cmul(e->zu, e->Yc));
}
/* unconjugated variable */
static inline ocplx d_unc(const eqshape *e, ocplx dK, ocplx dX, ocplx dY) {
    if (MUT(3)) return cadd(cmul(e->Z0, dK), cmul(csub(cmul(e->Z0, dX), cmul(e->U, dY)), e->Wc));
    return csub(cmul(e->Z0, dK), cmul(csub(cmul(e->Z0, dX), cmul(e->U, dY)), e->Wc));
}
/* conjugated variable */
static inline ocplx d_con(const eqshape *e, ocplx dK, ocplx dXc, ocplx dYc) {
    return csub(cmul(e->Z0, dK), cmul(e->W, csub(dXc, cmul(e->zu, dYc))));
}
static inline ocplx d_U(const eqshape *e) {
    return csub(cmul(e->Y, e->Wc), cdiv(cmul(cmul(e->W, e->zu), e->Yc), e->U));
}
static inline ocplx d_Z0(const eqshape *e, ocplx dXdZ0, ocplx dYdZ0, ocplx dXcdZ0, ocplx dYcdZ0) {
    ocplx dW = csub(cadd(e->X, cmul(e->Z0, dXdZ0)), cmul(e->U, dYdZ0));
    ocplx dWc = csub(csub(dXcdZ0, cdiv(e->Yc, e->U)), cmul(e->zu, dYcdZ0));
    return csub(csub(e->K, cmul(dW, e->Wc)), cmul(e->W, dWc));
}

/* III_j intermediates shared by JVAR and JPAR */
typedef struct {
    ocplx iFB, iHC, r1, r2, m, n, y1, y2;       /* unconjugated */
    ocplx iFBc, iHCc, r1c, r2c, mc, nc, y1c, y2c; /* conjugated */
} eq3aux;

static void build_I(const ocplx *Z, const ocplx *q, int j, eqshape *e) {
    ocplx A = Z[iA], Cv = Z[iC], H = Z[iH], Ac = Z[iAc], Cc = Z[iCc], Hc = Z[iHc];
    ocplx Pj = q[j], P0 = q[0], Pcj = q[8 + j], Pc0 = q[8], Q = q[15 + j], Qc = q[22 + j];
    e->Z0 = Z[21]; e->U = Z[13 + j]; e->zu = cdiv(e->Z0, e->U);
    ocplx CA = csub(Cv, A), CcAc = csub(Cc, Ac);
    e->X = csub(csub(cmul(Pj, e->Z0), A), cmul(CA, Q));
    e->Y = csub(cmul(P0, e->Z0), H);
    e->Xc = csub(csub(cmul(Pcj, e->Z0), Ac), cmul(CcAc, Qc));
    e->Yc = csub(cmul(Pc0, e->Z0), Hc);
    /* K = Ac C - Ac A - Hc C + A Cc - H Cc + H Hc + (C-A)(Cc-Ac) Q Qc */
    ocplx K = csub(cmul(Ac, Cv), cmul(Ac, A));
    K = csub(K, cmul(Hc, Cv));
    K = cadd(K, cmul(A, Cc));
    K = csub(K, cmul(H, Cc));
    K = cadd(K, cmul(H, Hc));
    if (!MUT(4)) K = cadd(K, cmul(cmul(CA, CcAc), cmul(Q, Qc)));
    e->K = K;
    shape_finish(e);
}
static void build_II(const ocplx *Z, const ocplx *q, int j, eqshape *e) {
    ocplx B = Z[iB], F = Z[iF], Bc = Z[iBc], Fc = Z[iFc];
    ocplx Pj = q[j], P0 = q[0], Pcj = q[8 + j], Pc0 = q[8];
    e->Z0 = Z[21]; e->U = Z[13 + j]; e->zu = cdiv(e->Z0, e->U);
    e->X = csub(cmul(Pj, e->Z0), B);
    e->Y = csub(cmul(P0, e->Z0), F);
    e->Xc = csub(cmul(Pcj, e->Z0), Bc);
    e->Yc = csub(cmul(Pc0, e->Z0), Fc);
    e->K = cmul(csub(B, F), csub(Bc, Fc));
    shape_finish(e);
}
static void build_III(const ocplx *Z, const ocplx *q, int j, eqshape *e, eq3aux *a) {
    ocplx A = Z[iA], B = Z[iB], Cv = Z[iC], D = Z[iD], F = Z[iF], G = Z[iG], H = Z[iH];
    ocplx Ac = Z[iAc], Bc = Z[iBc], Cc = Z[iCc], Dc = Z[iDc], Fc = Z[iFc], Gc = Z[iGc], Hc = Z[iHc];
    ocplx Pj = q[j], P0 = q[0], Pcj = q[8 + j], Pc0 = q[8], Q = q[15 + j], Qc = q[22 + j];
    ocplx one = C(1, 0);
    e->Z0 = Z[21]; e->U = Z[13 + j]; e->zu = cdiv(e->Z0, e->U);
    a->iFB = cdiv(one, csub(F, B)); a->iHC = cdiv(one, csub(H, Cv));
    a->r1 = cmul(csub(D, B), a->iFB); a->r2 = cmul(csub(G, Cv), a->iHC);
    a->m = csub(cmul(Pj, e->Z0), B);
    a->n = csub(csub(cmul(Pj, e->Z0), A), cmul(csub(Cv, A), Q));
    a->y1 = csub(cmul(P0, e->Z0), F); a->y2 = csub(cmul(P0, e->Z0), H);
    a->iFBc = cdiv(one, csub(Fc, Bc)); a->iHCc = cdiv(one, csub(Hc, Cc));
    a->r1c = cmul(csub(Dc, Bc), a->iFBc); a->r2c = cmul(csub(Gc, Cc), a->iHCc);
    a->mc = csub(cmul(Pcj, e->Z0), Bc);
    a->nc = csub(csub(cmul(Pcj, e->Z0), Ac), cmul(csub(Cc, Ac), Qc));
    a->y1c = csub(cmul(Pc0, e->Z0), Fc); a->y2c = csub(cmul(Pc0, e->Z0), Hc);
    /* X = (A-B) + (C-A)Q - r1 m + r2 n ; Y = -r1 y1 + r2 y2 */
    e->X = cadd(csub(cadd(csub(A, B), cmul(csub(Cv, A), Q)), cmul(a->r1, a->m)), cmul(a->r2, a->n));
    e->Y = csub(cmul(a->r2, a->y2), cmul(a->r1, a->y1));
    e->Xc = cadd(csub(cadd(csub(Ac, Bc), cmul(csub(Cc, Ac), Qc)), cmul(a->r1c, a->mc)),
                 cmul(a->r2c, a->nc));
    e->Yc = csub(cmul(a->r2c, a->y2c), cmul(a->r1c, a->y1c));
    e->K = cmul(csub(G, D), csub(Gc, Dc));
    shape_finish(e);
}

void oracle_jvar(const ocplx *Z, const ocplx *q, ocplx *J) {
    const ocplx one = C(1, 0), zero = C(0, 0);
    memset(J, 0, sizeof(ocplx) * ONV * (ONV + 1));
    for (int j = 1; j <= 7; ++j) {
        ocplx Pj = q[j], P0 = q[0], Pcj = q[8 + j], Pc0 = q[8], Q = q[15 + j], Qc = q[22 + j];
        /* I_j : nonzero columns A, C, H, Ac, Cc, Hc, U_j, Z0 */
        {
            eqshape e; build_I(Z, q, j, &e);
            ocplx *row = &J[(j - 1) * 22];
            ocplx A = Z[iA], Cv = Z[iC], H = Z[iH], Ac = Z[iAc], Cc = Z[iCc], Hc = Z[iHc];
            ocplx CA = csub(Cv, A), CcAc = csub(Cc, Ac), QQ = cmul(Q, Qc);
            ocplx omQQ = csub(one, QQ);
            /* dX/dA = Q - 1, dX/dC = -Q, dY/dH = -1 (and the same for the conjugates) */
            row[iA] = d_unc(&e, cmul(CcAc, omQQ), csub(Q, one), zero);
            row[iC] = d_unc(&e, cadd(csub(Ac, Hc), cmul(CcAc, QQ)), cneg(Q), zero);
            row[iH] = d_unc(&e, csub(Hc, Cc), zero, cneg(one));
            row[iAc] = d_con(&e, cmul(CA, omQQ), csub(Qc, one), zero);
            row[iCc] = d_con(&e, cadd(csub(A, H), cmul(CA, QQ)), cneg(Qc), zero);
            row[iHc] = d_con(&e, csub(H, Cv), zero, cneg(one));
            row[13 + j] = d_U(&e);
            row[21] = d_Z0(&e, Pj, P0, Pcj, Pc0);
        }
        /* II_j : nonzero columns B, F, Bc, Fc, U_j, Z0 */
        {
            eqshape e; build_II(Z, q, j, &e);
            ocplx *row = &J[(7 + j - 1) * 22];
            ocplx B = Z[iB], F = Z[iF], Bc = Z[iBc], Fc = Z[iFc];
            row[iB] = d_unc(&e, csub(Bc, Fc), cneg(one), zero);
            row[iF] = d_unc(&e, cneg(csub(Bc, Fc)), zero, cneg(one));
            row[iBc] = d_con(&e, csub(B, F), cneg(one), zero);
            row[iFc] = d_con(&e, cneg(csub(B, F)), zero, cneg(one));
            row[13 + j] = d_U(&e);
            row[21] = d_Z0(&e, Pj, P0, Pcj, Pc0);
        }
        /* III_j : all 14 linkage columns, U_j, Z0 */
        {
            eqshape e; eq3aux a; build_III(Z, q, j, &e, &a);
            ocplx *row = &J[(14 + j - 1) * 22];
            ocplx D = Z[iD], G = Z[iG], Dc = Z[iDc], Gc = Z[iGc];
            ocplx GD = csub(G, D), GDc = csub(Gc, Dc);
            /* unconjugated partials of X and Y */
            ocplx dXA = cmul(csub(one, Q), csub(one, a.r2));
            ocplx dXB = cmul(csub(a.r1, one), csub(one, cmul(a.m, a.iFB)));
            ocplx dYB = cneg(cmul(cmul(a.y1, csub(a.r1, one)), a.iFB));
            ocplx dXC = cmul(csub(one, a.r2), csub(Q, cmul(a.n, a.iHC)));
            ocplx dYC = cmul(cmul(a.y2, csub(a.r2, one)), a.iHC);
            ocplx dXD = cneg(cmul(a.m, a.iFB));
            ocplx dYD = cneg(cmul(a.y1, a.iFB));
            ocplx dXF = cmul(cmul(a.m, a.r1), a.iFB);
            ocplx dYF = cmul(a.r1, cadd(one, cmul(a.y1, a.iFB)));
            ocplx dXG = cmul(a.n, a.iHC);
            ocplx dYG = cmul(a.y2, a.iHC);
            ocplx dXH = cneg(cmul(cmul(a.n, a.r2), a.iHC));
            ocplx dYH = cneg(cmul(a.r2, cadd(one, cmul(a.y2, a.iHC))));
            row[iA] = d_unc(&e, zero, dXA, zero);
            row[iB] = d_unc(&e, zero, dXB, dYB);
            row[iC] = d_unc(&e, zero, dXC, dYC);
            row[iD] = d_unc(&e, cneg(GDc), dXD, dYD);
            row[iF] = d_unc(&e, zero, dXF, dYF);
            row[iG] = d_unc(&e, GDc, dXG, dYG);
            row[iH] = d_unc(&e, zero, dXH, dYH);
            /* conjugated partials */
            ocplx dXAc = cmul(csub(one, Qc), csub(one, a.r2c));
            ocplx dXBc = cmul(csub(a.r1c, one), csub(one, cmul(a.mc, a.iFBc)));
            ocplx dYBc = cneg(cmul(cmul(a.y1c, csub(a.r1c, one)), a.iFBc));
            ocplx dXCc = cmul(csub(one, a.r2c), csub(Qc, cmul(a.nc, a.iHCc)));
            ocplx dYCc = cmul(cmul(a.y2c, csub(a.r2c, one)), a.iHCc);
            ocplx dXDc = cneg(cmul(a.mc, a.iFBc));
            ocplx dYDc = cneg(cmul(a.y1c, a.iFBc));
            ocplx dXFc = cmul(cmul(a.mc, a.r1c), a.iFBc);
            ocplx dYFc = cmul(a.r1c, cadd(one, cmul(a.y1c, a.iFBc)));
            ocplx dXGc = cmul(a.nc, a.iHCc);
            ocplx dYGc = cmul(a.y2c, a.iHCc);
            ocplx dXHc = cneg(cmul(cmul(a.nc, a.r2c), a.iHCc));
            ocplx dYHc = cneg(cmul(a.r2c, cadd(one, cmul(a.y2c, a.iHCc))));
            row[iAc] = d_con(&e, zero, dXAc, zero);
            row[iBc] = d_con(&e, zero, dXBc, dYBc);
            row[iCc] = d_con(&e, zero, dXCc, dYCc);
            row[iDc] = d_con(&e, cneg(GD), dXDc, dYDc);
            row[iFc] = d_con(&e, zero, dXFc, dYFc);
            row[iGc] = d_con(&e, GD, dXGc, dYGc);
            row[iHc] = d_con(&e, zero, dXHc, dYHc);
            row[13 + j] = d_U(&e);
            ocplx dr = csub(a.r2, a.r1), drc = csub(a.r2c, a.r1c);
            row[21] = d_Z0(&e, cmul(Pj, dr), cmul(P0, dr), cmul(Pcj, drc), cmul(Pc0, drc));
        }
    }
}

/* dF/dq (htpykinematics.cu:1132-1384): per pose only P0, P_j, Pc0, Pc_j, Q_j, Qc_j enter. */
void oracle_jpar(const ocplx *Z, const ocplx *q, ocplx *J) {
    memset(J, 0, sizeof(ocplx) * ONV * ONP);
    const ocplx one = C(1, 0);
    for (int j = 1; j <= 7; ++j) {
        ocplx Q = q[15 + j], Qc = q[22 + j];
        {
            eqshape e; build_I(Z, q, j, &e);
            ocplx *row = &J[(j - 1) * 30];
            ocplx CA = csub(Z[iC], Z[iA]), CcAc = csub(Z[iCc], Z[iAc]);
            ocplx CACc = cmul(CA, CcAc);
            row[0] = cmul(cmul(e.U, e.Z0), e.Wc);                       /* P0 */
            row[j] = cneg(cmul(cmul(e.Z0, e.Z0), e.Wc));                /* P_j */
            row[8] = cmul(cmul(e.W, e.zu), e.Z0);                       /* Pc0 */
            row[8 + j] = cneg(cmul(e.W, e.Z0));                         /* Pc_j */
            row[15 + j] = cadd(cmul(cmul(e.Z0, CACc), Qc), cmul(cmul(e.Z0, CA), e.Wc)); /* Q_j */
            row[22 + j] = cadd(cmul(cmul(e.Z0, CACc), Q), cmul(e.W, CcAc));             /* Qc_j */
        }
        {
            eqshape e; build_II(Z, q, j, &e);
            ocplx *row = &J[(7 + j - 1) * 30];
            row[0] = cmul(cmul(e.U, e.Z0), e.Wc);
            row[j] = cneg(cmul(cmul(e.Z0, e.Z0), e.Wc));
            row[MUT(5) ? 8 + j : 8] = cmul(cmul(e.W, e.zu), e.Z0);
            row[MUT(5) ? 8 : 8 + j] = cneg(cmul(e.W, e.Z0));
        }
        {
            eqshape e; eq3aux a; build_III(Z, q, j, &e, &a);
            ocplx *row = &J[(14 + j - 1) * 30];
            ocplx dr = csub(a.r2, a.r1), drc = csub(a.r2c, a.r1c);
            ocplx CA = csub(Z[iC], Z[iA]), CcAc = csub(Z[iCc], Z[iAc]);
            row[0] = cmul(cmul(e.U, cmul(e.Z0, dr)), e.Wc);
            row[j] = cneg(cmul(cmul(e.Z0, cmul(e.Z0, dr)), e.Wc));
            row[15 + j] = cneg(cmul(cmul(e.Z0, cmul(CA, csub(one, a.r2))), e.Wc));
            row[8] = cmul(cmul(e.W, e.zu), cmul(e.Z0, drc));
            row[8 + j] = cneg(cmul(e.W, cmul(e.Z0, drc)));
            row[22 + j] = cneg(cmul(e.W, cmul(CcAc, csub(one, a.r2c))));
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* L2: homotopy support (htpysupport.cu)                                                      */
/* ------------------------------------------------------------------------------------------ */

/* SwitchPatch (htpysupport.cu:9-34) */
void oracle_switch_patch(const ocplx *unew, const ocplx *Yold, ocplx *Ynew) {
    ocplx ut = unew[ONV + 1], t = Yold[ONV + 1];
    ocplx den = C(0, 0);
    for (int i = 0; i < ONV + 1; ++i) den = cadd(den, cmul(unew[i], Yold[i]));
    ocplx scalar = cdiv(csub(C(1, 0), cmul(ut, t)), den);
    ocplx tmp[ONY];
    for (int i = 0; i < ONV + 1; ++i) tmp[i] = cmul(scalar, Yold[i]);
    tmp[ONV + 1] = MUT(7) ? cmul(scalar, t) : t;
    memcpy(Ynew, tmp, sizeof(tmp));
}

/* q (htpysupport.cu:37-48): q(t) = (gamma (1-t) s + t f) / (gamma (1-t) + t) */
void oracle_q(double t, const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *out) {
    for (int i = 0; i < ONP; ++i) {
        out[i] = cdiv(cadd(cmul(cmul(*gamma, C(1. - t, 0)), spm[i]), cmul(C(t, 0), fpm[i])),
                      MUT(6) ? cadd(cmul(*gamma, C(t, 0)), C(1. - t, 0)) : cadd(cmul(*gamma, C(1. - t, 0)), C(t, 0)));
    }
}

/* dqdt (htpysupport.cu:51-62): gamma (f - s) / (gamma (1-t) + t)^2 */
void oracle_dqdt(double t, const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *out) {
    for (int i = 0; i < ONP; ++i) {
        out[i] = cdiv(cmul(csub(fpm[i], spm[i]), *gamma),
                      cpow2(cadd(cmul(*gamma, C(1. - t, 0)), C(t, 0))));
    }
}

/* H (htpysupport.cu:64-86): [F(Z; q(t)); u.Y - 1] */
void oracle_H(const ocplx *Y, const ocplx *u, const ocplx *spm, const ocplx *fpm,
              const ocplx *gamma, ocplx *out) {
    ocplx params[ONP];
    double t = Y[ONV + 1].re;
    oracle_q(t, spm, fpm, gamma, params);
    oracle_fsys(Y, params, out);
    out[ONV] = csub(vdot(u, Y, ONY), C(1, 0));
}

/* JH (htpysupport.cu:89-144): 22 x 23, rows 0..20 = [dF/dz, dF/dZ0, dF/dq . dq/dt], row 21 = u.
 * The reference assembles in place inside the output buffer; here separate buffers are used
 * (same values). */
void oracle_JH(const ocplx *Y, const ocplx *u, const ocplx *spm, const ocplx *fpm,
               const ocplx *gamma, ocplx *out) {
    ocplx params[ONP], dparams[ONP];
    ocplx A[ONV * (ONV + 1)], dFdq[ONV * ONP];
    double t = Y[ONV + 1].re;
    oracle_q(t, spm, fpm, gamma, params);
    oracle_jvar(Y, params, A);
    oracle_jpar(Y, params, dFdq);
    oracle_dqdt(t, spm, fpm, gamma, dparams);
    for (int i = 0; i < ONV; ++i) {
        /* Matr::mul (classes.cuh:190-199) */
        ocplx b = C(0, 0);
        for (int k = 0; k < ONP; ++k) b = cadd(b, cmul(dFdq[i * ONP + k], dparams[k]));
        for (int k = 0; k < ONV + 1; ++k) out[i * ONY + k] = A[i * (ONV + 1) + k];
        out[i * ONY + ONV + 1] = b;
    }
    for (int k = 0; k < ONY; ++k) out[ONV * ONY + k] = u[k];
}

/* LinearSolve (htpysupport.cu:146-232): 23 x 23 LU with partial pivoting on |a|^2 (first
 * maximum wins), row-pointer swaps, then back substitution. */
void oracle_linear_solve(ocplx *mat, ocplx *b, ocplx *x) {
    const int n = ONY;
    ocplx *a[ONY];
    for (int i = 0; i < n; ++i) a[i] = &mat[i * n];
    for (int i = 0; i < n; ++i) {
        double maxAbs = 0.0, val;
        int maxRow = i;
        for (int k = i; k < n; ++k) {
            val = a[k][i].re * a[k][i].re + a[k][i].im * a[k][i].im;
            if (MUT(16) ? val < maxAbs : val > maxAbs) {
                maxAbs = val;
                maxRow = k;
            }
        }
        if (maxRow != i) {
            ocplx *p = a[i]; a[i] = a[maxRow]; a[maxRow] = p;
            if (!MUT(8)) { ocplx tb = b[i]; b[i] = b[maxRow]; b[maxRow] = tb; }
        }
        for (int j = i + 1; j < n; ++j) {
            a[j][i] = cdiv(a[j][i], a[i][i]);
            for (int k = i + 1; k < n; ++k) a[j][k] = csub(a[j][k], cmul(a[i][k], a[j][i]));
            b[j] = csub(b[j], cmul(b[i], a[j][i]));
        }
    }
    for (int i = n - 1; i > -1; --i) {
        ocplx sum = C(0, 0);
        for (int k = i + 1; k < n; ++k) sum = cfma(a[i][k], x[k], sum);
        x[i] = cmul(cdiv(C(1, 0), a[i][i]), csub(b[i], sum));
    }
}

/* LinearSolveNewton (htpysupport.cu:302-400): pivots and eliminates over rows 0..21 only, all
 * 23 columns carried, x_22 := tdiff, back substitution over rows 21..0. */
void oracle_linear_solve_newton(ocplx *mat, ocplx *b, double tdiff, ocplx *x) {
    const int n = ONY;
    ocplx *a[ONY];
    for (int i = 0; i < n; ++i) a[i] = &mat[i * n];
    for (int i = 0; i < n; ++i) {
        double maxAbs = 0.0, val;
        int maxRow = i;
        for (int k = i; k < n - 1; ++k) {
            if ((val = a[k][i].re * a[k][i].re + a[k][i].im * a[k][i].im) > maxAbs) {
                maxAbs = val;
                maxRow = k;
            }
        }
        if (maxRow != i) {
            ocplx *p = a[i]; a[i] = a[maxRow]; a[maxRow] = p;
            ocplx tb = b[i]; b[i] = b[maxRow]; b[maxRow] = tb;
        }
        for (int j = i + 1; j < n - 1; ++j) {
            a[j][i] = cdiv(a[j][i], a[i][i]);
            for (int k = i + 1; k < n; ++k) a[j][k] = csub(a[j][k], cmul(a[i][k], a[j][i]));
            b[j] = csub(b[j], cmul(b[i], a[j][i]));
        }
    }
    x[n - 1] = C(MUT(20) ? -tdiff : tdiff, 0.);
    for (int i = n - 2; i > -1; --i) {
        ocplx sum = C(0, 0);
        for (int k = i + 1; k < n; ++k) sum = cfma(a[i][k], x[k], sum);
        x[i] = cmul(cdiv(C(1, 0), a[i][i]), csub(b[i], sum));
    }
}

/* Row scaling shared by IVP and Newton: rows 0..21 of the 23 x 23 matrix are scaled by
 * 1/max(|re|,|im|) over the row (htpysupport.cu:260-277, :446-471).  Returns the scales. */
static void scale_rows(ocplx *M, double *scales) {
    for (int i = 0; i < ONV + 1; ++i) {
        double maxVal = 0.0;
        for (int j = 0; j < ONY; ++j) {
            if (fabs(M[i * ONY + j].re) > maxVal) maxVal = fabs(M[i * ONY + j].re);
            if (fabs(M[i * ONY + j].im) > maxVal) maxVal = fabs(M[i * ONY + j].im);
        }
        double s = 1 / maxVal;
        scales[i] = s;
        for (int j = 0; j < ONY; ++j) {
            M[i * ONY + j].re = s * M[i * ONY + j].re;
            M[i * ONY + j].im = s * M[i * ONY + j].im;
        }
    }
}

/* IVP (htpysupport.cu:237-299) */
void oracle_ivp(int tstep, const ocplx *Yc, const ocplx *Vc, const ocplx *u, const ocplx *spm,
                const ocplx *fpm, const ocplx *gamma, ocplx *dYds) {
    ocplx M[ONY * ONY], b[ONY], x[ONY];
    double scales[ONV + 1];
    oracle_JH(Yc, u, spm, fpm, gamma, M);
    for (int j = 0; j < ONY; ++j) M[(ONV + 1) * ONY + j] = Vc[j];
    scale_rows(M, scales);
    memset(b, 0, sizeof(b));
    b[ONV + 1] = C(1.0, 0);
    oracle_linear_solve(M, b, x);
    ocplx sc;
    if (!tstep) {
        double theta = atan2(x[ONV + 1].im, x[ONV + 1].re);
        sc = cmul(C(cos(theta), -sin(theta)), C(MUT(19) ? 1.0 : 1.0 / vnorm(x, ONY), 0));
    } else {
        sc = cdiv(C(1.0, 0), x[ONV + 1]);
    }
    for (int i = 0; i < ONY; ++i) dYds[i] = cmul(sc, x[i]);
    dYds[ONV + 1].im = 0.;
}

/* Newton (htpysupport.cu:407-494).  Returns the number of iterations performed. */
int oracle_newton(double tgoal, const ocplx *Vn, const ocplx *u, const ocplx *spm, const ocplx *fpm,
                  const ocplx *gamma, int max_its, double tol, int residual_type, ocplx *YN,
                  double *residual) {
    int i = 0;
    *residual = 100.0;
    while (i < max_its && *residual > tol) {
        ocplx M[ONY * ONY], Hv[ONV + 1], bvec[ONV + 1], Yd[ONY];
        double scales[ONV + 1];
        oracle_JH(YN, u, spm, fpm, gamma, M);
        for (int j = 0; j < ONY; ++j) M[(ONV + 1) * ONY + j] = Vn[j];
        oracle_H(YN, u, spm, fpm, gamma, Hv);
        scale_rows(M, scales);
        for (int j = 0; j < ONV + 1; ++j) {
            Hv[j].re = scales[j] * Hv[j].re;
            Hv[j].im = scales[j] * Hv[j].im;
        }
        double tdiff = tgoal - YN[ONV + 1].re;
        for (int j = 0; j < ONV + 1; ++j) bvec[j] = crs(-1.0, Hv[j]);
        oracle_linear_solve_newton(M, bvec, tdiff, Yd);
        for (int j = 0; j < ONY; ++j) YN[j] = MUT(9) ? csub(YN[j], Yd[j]) : cadd(YN[j], Yd[j]);
        if (residual_type == 0) {
            oracle_H(YN, u, spm, fpm, gamma, Hv);
            *residual = vnorm(Hv, ONV + 1);
        } else {
            *residual = vnorm(Yd, ONY);
        }
        i++;
    }
    return i;
}

/* ------------------------------------------------------------------------------------------ */
/* L3: Dormand-Prince RK6(5)8M step (rungekutta.cu:3-124)                                     */
/* ------------------------------------------------------------------------------------------ */

/* y = base + (c0 k0 + (c1 k1 + ( ... + c_{n-1} k_{n-1}))), nested as in rungekutta.cu:43-122 */
static void rk_combo(const ocplx *base, int n, const double *c, const ocplx *const *ks, ocplx *y) {
    for (int i = 0; i < ONY; ++i) {
        ocplx acc = crs(c[n - 1], ks[n - 1][i]);
        for (int m = n - 2; m >= 0; --m) acc = cadd(crs(c[m], ks[m][i]), acc);
        y[i] = base ? cadd(base[i], acc) : acc;
    }
}

void oracle_rk65(double s, double tol, int tstep, const ocplx *Yc, const ocplx *Vc, const ocplx *u,
                 const ocplx *spm, const ocplx *fpm, const ocplx *gamma, ocplx *Yn, ocplx *Vn,
                 double *err, double *scale) {
    ocplx k[8][ONY], Yarg[ONY], LTE[ONY];
    oracle_ivp(tstep, Yc, Vc, u, spm, fpm, gamma, k[0]);
    {
        double c[1] = {1. / 10. * s};
        const ocplx *ks[1] = {k[0]};
        rk_combo(Yc, 1, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[0], u, spm, fpm, gamma, k[1]);
    }
    {
        double c[2] = {-2. / 81. * s, 20. / 81. * s};
        const ocplx *ks[2] = {k[0], k[1]};
        rk_combo(Yc, 2, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[1], u, spm, fpm, gamma, k[2]);
    }
    {
        double c[3] = {615. / 1372. * s, -270. / 343. * s, 1053. / 1372. * s};
        const ocplx *ks[3] = {k[0], k[1], k[2]};
        rk_combo(Yc, 3, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[2], u, spm, fpm, gamma, k[3]);
    }
    {
        double c[4] = {3243. / 5500. * s, -54. / 55. * s, 50949. / 71500. * s, 4998. / 17875. * s};
        const ocplx *ks[4] = {k[0], k[1], k[2], k[3]};
        rk_combo(Yc, 4, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[3], u, spm, fpm, gamma, k[4]);
    }
    {
        double c[5] = {-26492. / 37125. * s, 72. / 55. * s, 2808. / 23375. * s,
                       -24206. / 37125. * s, 338. / 459. * s};
        const ocplx *ks[5] = {k[0], k[1], k[2], k[3], k[4]};
        rk_combo(Yc, 5, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[4], u, spm, fpm, gamma, k[5]);
    }
    {
        double c[6] = {5561. / 2376. * s, -35. / 11. * s, -24117. / 31603. * s,
                       899983. / 200772. * s, -5225. / 1836. * s, 3925. / 4056. * s};
        const ocplx *ks[6] = {k[0], k[1], k[2], k[3], k[4], k[5]};
        rk_combo(Yc, 6, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[5], u, spm, fpm, gamma, k[6]);
    }
    {
        double c[6] = {465467. / 266112. * s, -2945. / 1232. * s, -5610201. / 14158144. * s,
                       10513573. / 3212352. * s, -424325. / 205632. * s, 376225. / 454272. * s};
        const ocplx *ks[6] = {k[0], k[1], k[2], k[3], k[4], k[5]};
        rk_combo(Yc, 6, c, ks, Yarg);
        oracle_ivp(tstep, Yarg, k[6], u, spm, fpm, gamma, k[7]);
    }
    {
        double c[7] = {-13. / 2400. * s, 19683. / 618800. * s, -2401. / 31200. * s, 65. / 816. * s,
                       -15. / 416. * s, -521. / 5600. * s, 1. / 10. * s};
        const ocplx *ks[7] = {k[0], k[2], k[3], k[4], k[5], k[6], k[7]};
        rk_combo(NULL, 7, c, ks, LTE);
    }
    *err = vnorm(LTE, ONY);
    if (*err > 10e-17) *scale = 0.8 * pow((tol / (*err)), MUT(11) ? 1. / 5. : 1. / 6.);
    else *scale = 5.0;
    if (*scale > 5) *scale = 5;
    {
        double c[7] = {61. / 864. * s, 98415. / 321776. * s, 16807. / 146016. * s,
                       1375. / 7344. * s, 1375. / 5408. * s, -37. / 1120. * s, 1. / 10. * s};
        if (MUT(10)) { const double t = c[3]; c[3] = c[4]; c[4] = t; }
        const ocplx *ks[7] = {k[0], k[2], k[3], k[4], k[5], k[6], k[7]};
        rk_combo(Yc, 7, c, ks, Yn);
    }
    memcpy(Vn, k[7], sizeof(ocplx) * ONY);
}

/* ------------------------------------------------------------------------------------------ */
/* L4: TrackPath (htpytrack.cu:33-220)                                                         */
/* ------------------------------------------------------------------------------------------ */

static void draw_patch(orng *g, const otrk_opts *o, ocplx *u) {
    for (int i = 0; i < ONY; ++i) {
        u[i].re = (o->patch_hi - o->patch_lo) * rng_next(g) + o->patch_lo;
        u[i].im = (o->patch_hi - o->patch_lo) * rng_next(g) + o->patch_lo;
    }
}

int oracle_track_path(const otrk_opts *o, const ocplx *spt, const ocplx *spm, const ocplx *fpm,
                      const ocplx *gamma, uint64_t seed, uint64_t path_id, uint64_t first_draw,
                      ocplx *Yfinal, otrk_stats *st) {
    orng g = {seed, path_id, first_draw};
    ocplx Vc[ONY], Yn[ONY], Vn[ONY], YN[ONY], u[ONY];
    ocplx *Yc = Yfinal; /* Yc aliases Yfinal, htpytrack.cu:46 */
    otrk_stats S;
    memset(&S, 0, sizeof(S));
    int code = 0;
    for (int i = 0; i < ONV; ++i) Yc[i] = spt[i];
    Yc[ONV] = C(1, 0);
    Yc[ONV + 1] = C(0, 0);
    draw_patch(&g, o, u);
    oracle_switch_patch(u, Yc, Yc);
    draw_patch(&g, o, Vc);
    oracle_ivp(0, Yc, Vc, u, spm, fpm, gamma, Vc);
    S.ivp_calls++;

    double s = o->sinit, err = 0, scale = 1, residual = 100;
    int tstepQ = 0;
    int streak = o->caution_steps - 1;
    int step = 0, ministep = 0;

    for (;;) { /* main loop */
        if (step == o->max_steps) { code = 1; goto done; }
        for (;;) { /* correction loop */
            for (;;) { /* final steps loop */
                for (;;) { /* predictor loop */
                    oracle_rk65(s, o->rk_tol, tstepQ, Yc, Vc, u, spm, fpm, gamma, Yn, Vn, &err,
                                &scale);
                    S.rk_attempts++;
                    S.ivp_calls += 8;
                    if ((MUT(25) ? ++ministep : ministep++) == o->max_ministeps) { code = 2; goto done; }
                    if (err <= o->rk_tol) break;
                    if (s <= o->smin) { code = 3; goto done; }
                    S.rejections++;
                    s *= scale;
                    if (s < o->smin) s = o->smin * 0.9;
                    if (!MUT(23)) streak = 0;
                }
                if (Yn[ONV + 1].re <= 1. - o->nudge_trigger) break;
                if (fabs(1. - Yn[ONV + 1].re) < o->sharp_trigger) break;
                tstepQ = 1;
                s = 1 - (MUT(26) ? Yn : Yc)[ONV + 1].re;
            }
            if (fabs(1. - Yn[ONV + 1].re) < o->sharp_trigger) break;
            int i = 0;
            while (i < o->newt_attempts) {
                memcpy(YN, Yn, sizeof(YN));
                S.newton_its += oracle_newton(Yn[ONV + 1].re, Vn, u, spm, fpm, gamma, o->newt_its,
                                              o->newt_tol, o->newt_residual, YN, &residual);
                S.newton_calls++;
                if (ministep++ == o->max_ministeps) { code = 2; goto done; }
                if (residual < o->newt_tol) break;
                if (!MUT(24)) {
                    draw_patch(&g, o, u);
                    oracle_switch_patch(u, Yn, Yn);
                }
                S.redraws++;
                i++;
            }
            if (residual < o->newt_tol) break;
            if (s <= o->smin) { code = 4; goto done; }
            s *= 0.5;
            if (s < o->smin) s = o->smin * 0.9;
            streak = 0;
        }
        if (fabs(1. - Yn[ONV + 1].re) < o->sharp_trigger) break;
        memcpy(Yc, YN, sizeof(YN));
        memcpy(Vc, Vn, sizeof(Vn));
        streak++;
        if (!MUT(22) && streak < o->caution_steps && scale > 1.) scale = 1.;
        s *= scale;
        if (s > o->smax) s = o->smax;
        else if (s < o->smin) s = o->smin * 0.9;
        step++;
    }
    /* sharpening (htpytrack.cu:194-216) */
    {
        int i = 0;
        while (i < o->sharp_attempts) {
            memcpy(Yfinal, Yn, sizeof(Yn));
            S.newton_its += oracle_newton(1., Vn, u, spm, fpm, gamma, o->sharp_its, o->sharp_tol,
                                          o->sharp_residual, Yfinal, &residual);
            S.newton_calls++;
            if (ministep++ == o->max_ministeps) { code = 2; goto done; }
            if (residual < o->sharp_tol) break;
            draw_patch(&g, o, u);
            S.redraws++;
            oracle_switch_patch(u, Yn, Yn);
            i++;
        }
        if (residual >= o->sharp_tol) { code = 5; goto done; }
    }
done:
    S.steps = step;
    S.ministeps = ministep;
    S.end_code = code;
    if (st) *st = S;
    return code;
}

/* ------------------------------------------------------------------------------------------ */
/* L5: start system and cognates (frgkinematics.cu)                                            */
/* ------------------------------------------------------------------------------------------ */

static ocplx draw_start_value(orng *g, const otrk_opts *o) {
    double a = rng_next(g), b = rng_next(g);
    if (o->start_mode == 1) {
        /* standard complex Gaussian (E|z|^2 = 1) by Box-Muller on the same two draws */
        double r = sqrt(-log(a)), th = 6.283185307179586 * b;
        return C(r * cos(th), r * sin(th));
    }
    return C((o->start_hi - o->start_lo) * a + o->start_lo, (o->start_hi - o->start_lo) * b + o->start_lo);
}

/* GenerateStart (frgkinematics.cu:13-136) */
void oracle_generate_start(const otrk_opts *o, uint64_t seed, uint64_t path_id, ocplx *z, ocplx *q) {
    orng g = {seed, path_id, 0};
    const ocplx one = C(1, 0), zero = C(0, 0);
    ocplx *P = &q[0], *Pc = &q[8], *Q = &q[15], *Qc = &q[22], *U = &z[13];
    for (int i = 0; i < 14; ++i) z[i] = draw_start_value(&g, o);
    P[0] = draw_start_value(&g, o);
    Pc[0] = draw_start_value(&g, o);
    ocplx *RUc = P; /* shares space with P (frgkinematics.cu:51) */
    for (int j = 1; j <= 7; ++j) RUc[j] = draw_start_value(&g, o);
    const ocplx A = z[iA], B = z[iB], Cv = z[iC], D = z[iD], F = z[iF], G = z[iG], H = z[iH];
    const ocplx Ac = z[iAc], Bc = z[iBc], Cc = z[iCc], Dc = z[iDc], Fc = z[iFc], Gc = z[iGc],
                Hc = z[iHc];
    for (int j = 1; j <= 7; ++j) {
        ocplx R = RUc[j], iR = cdiv(one, RUc[j]);
        ocplx A_quad = cmul(csub(G, D), csub(cmul(iR, csub(Hc, Gc)), csub(Hc, Fc)));
        ocplx Ac_quad = cmul(csub(Gc, Dc), csub(cmul(R, csub(H, G)), csub(H, F)));
        ocplx B_quad = cadd(
            csub(csub(zero, cmul(R, cmul(csub(H, G), csub(Hc, Fc)))),
                 cmul(iR, cmul(csub(Hc, Gc), csub(H, F)))),
            cadd(cmul(csub(G, D), csub(Gc, Dc)),
                 cadd(cmul(csub(H, G), csub(Hc, Gc)),
                      csub(cmul(csub(H, F), csub(Hc, Fc)), cmul(csub(F, D), csub(Fc, Dc))))));
        double pm = 1.; if (rng_next(&g) < 0.5) pm = -1.;
        ocplx TUc = cmul(cdiv(one, cmul(C(2, 0), A_quad)),
                         cadd(csub(zero, B_quad),
                              cmul(C(pm, 0), csqrt_ref(csub(cmul(B_quad, B_quad),
                                                            cmul(C(4, 0), cmul(A_quad, Ac_quad)))))));
        ocplx SUc = cmul(cdiv(one, csub(F, D)),
                         cadd(cmul(TUc, csub(G, D)), csub(cmul(R, csub(H, G)), csub(H, F))));
        ocplx iSUc = cdiv(one, SUc);
        A_quad = cmul(csub(Ac, Bc), csub(csub(cmul(R, csub(H, Cv)), cmul(SUc, csub(F, B))), csub(H, F)));
        Ac_quad = cmul(csub(A, B), csub(csub(cmul(iR, csub(Hc, Cc)), cmul(iSUc, csub(Fc, Bc))),
                                        csub(Hc, Fc)));
        B_quad = cadd(cmul(csub(csub(cmul(R, csub(H, Cv)), cmul(SUc, csub(F, B))), csub(H, F)),
                           csub(csub(cmul(iR, csub(Hc, Cc)), cmul(iSUc, csub(Fc, Bc))), csub(Hc, Fc))),
                      csub(cmul(csub(A, B), csub(Ac, Bc)), cmul(csub(Cv, A), csub(Cc, Ac))));
        pm = 1.; if (rng_next(&g) < 0.5) pm = -1.;
        U[j] = cmul(cdiv(one, cmul(C(2, 0), A_quad)),
                    cadd(csub(zero, B_quad),
                         cmul(C(pm, 0), csqrt_ref(csub(cmul(B_quad, B_quad),
                                                       cmul(C(4, 0), cmul(A_quad, Ac_quad)))))));
        ocplx QUc = cmul(cdiv(C(-1, 0), csub(Cv, A)),
                         cadd(cmul(cdiv(one, U[j]), csub(A, B)),
                              csub(csub(cmul(R, csub(H, Cv)), cmul(SUc, csub(F, B))), csub(H, F))));
        Q[j] = cmul(QUc, U[j]);
        Qc[j] = MUT(15) ? Q[j] : cdiv(one, Q[j]);
        ocplx S = cmul(SUc, U[j]);
        P[j] = cadd(B, cadd(cmul(S, csub(F, B)), cmul(U[j], csub(P[0], F))));
        Pc[j] = cadd(Bc, cadd(cmul(cdiv(one, S), csub(Fc, Bc)), cmul(cdiv(one, U[j]), csub(Pc[0], Fc))));
    }
}

/* Cognate12 (frgkinematics.cu:185-224): the Stephenson II 1/2 cognate. */
void oracle_cognate12(const ocplx *orig, ocplx P, ocplx *cogn) {
    ocplx A = orig[0], B = orig[1], Cv = orig[2], D = orig[3], F = orig[4], G = orig[5], H = orig[6];
    ocplx scalar = cdiv(csub(B, F), csub(D, F));
    ocplx Dp = cadd(cmul(scalar, csub(D, H)), H);
    ocplx Gp = cadd(cmul(scalar, csub(G, H)), H);
    scalar = cdiv(csub(P, F), csub(H, F));
    cogn[0] = cadd(cmul(scalar, csub(A, B)), B);
    cogn[2] = cadd(cmul(scalar, csub(Cv, B)), B);
    cogn[3] = cadd(cmul(scalar, csub(Dp, B)), B);
    cogn[4] = cadd(B, MUT(12) ? csub(F, P) : csub(P, F));
    cogn[5] = cadd(cmul(scalar, csub(Gp, B)), B);
    cogn[6] = cadd(cmul(scalar, csub(H, B)), B);
    cogn[1] = B;
}

/* CognateSolns (frgkinematics.cu:227-285): out[0] = whichever of {z, its 1/2 cognate} has the
 * smaller Re(z_0); out[1] = the other one. */
void oracle_cognate_solns(const ocplx *z, const ocplx *q, ocplx *out) {
    ocplx c0[ONV], c1[ONV];
    memcpy(c0, z, sizeof(c0));
    ocplx B = z[iB], F = z[iF];
    const ocplx *P = &q[0], *Pc = &q[8];
    for (int i = 1; i <= 7; ++i)
        c1[13 + i] = cdiv(csub(csub(P[i], B), cmul(z[13 + i], csub(P[0], F))), csub(F, B));
    oracle_cognate12(&z[0], P[0], &c1[0]);
    oracle_cognate12(&z[7], Pc[0], &c1[7]);
    if (c1[0].re < c0[0].re) {
        memcpy(out, c1, sizeof(c1));
        memcpy(out + ONV, c0, sizeof(c0));
    } else {
        memcpy(out, c0, sizeof(c0));
        memcpy(out + ONV, c1, sizeof(c1));
    }
}

/* ------------------------------------------------------------------------------------------ */
/* L6: the Homotopy kernel body (frg.cu:35-69)                                                 */
/* ------------------------------------------------------------------------------------------ */

/* A refinement step longer than this fraction of max(1, |Y|) is discarded and the refinement
 * ends: the point is not in Newton's quadratic basin. */
#define POLISH_CAP 1e-6

/* Final FP64 refinement of a code-0 root (not in the reference; the final root refinement of the
 * tracker's output): polish_its Newton steps at t = 1 in the dehomogenising patch e_21 that
 * SwitchPatch has just applied (frg.cu:44-46,60), with Z0 = 1 exactly, so the root is refined in
 * the coordinates it is reported in.  Each step is iterative refinement with an extended
 * residual: the Jacobian, the row scaling and the solve are the double-precision Newton step of
 * htpysupport.cu:430-480 (JH with start = target so that q(1) = fpm exactly, LinearSolveNewton
 * with dt = 0), while F is evaluated in binary128 and rounded, so the iteration converges to the
 * double nearest the root even when the root is ill-conditioned.  A step is taken only while it
 * is no longer than POLISH_CAP * max(1, |Y|). */
static void polish_root(const otrk_opts *o, const ocplx *u_dehom, const ocplx *fpm, const ocplx *gamma, ocplx *Y,
                        otrk_stats *st) {
    Y[ONV] = C(1., 0);
    for (int k = 0; k < o->polish_its; ++k) {
        ocplx M[ONY * ONY], Hv[ONV + 1], b[ONV + 1], d[ONY];
        double scales[ONV + 1];
        oracle_JH(Y, u_dehom, fpm, fpm, gamma, M);
        for (int j = 0; j < ONY; ++j) M[(ONV + 1) * ONY + j] = C(0, 0); /* unused row */
        fsys_quad(Y, fpm, Hv);
        Hv[ONV] = csub(vdot(u_dehom, Y, ONY), C(1, 0));
        scale_rows(M, scales);
        for (int j = 0; j < ONV + 1; ++j) b[j] = C(-scales[j] * Hv[j].re, -scales[j] * Hv[j].im);
        oracle_linear_solve_newton(M, b, 0.0, d);
        const double r = vnorm(d, ONY);
        if (!(r <= POLISH_CAP * fmax(1.0, vnorm(Y, ONY)))) break;
        for (int j = 0; j < ONY; ++j) Y[j] = cadd(Y[j], d[j]);
        if (st) st->newton_its++;
    }
}

int oracle_refine_root(const otrk_opts *o, const ocplx *fpm, ocplx gamma, ocplx *Y) {
    ocplx u_dehom[ONY];
    memset(u_dehom, 0, sizeof(u_dehom));
    u_dehom[ONV] = C(1., 0);
    otrk_stats st;
    memset(&st, 0, sizeof(st));
    polish_root(o, u_dehom, fpm, &gamma, Y, &st);
    return st.newton_its;
}

/* Dehomogenise (frg.cu:60), polish code-0 roots, take the canonical cognate (frg.cu:63-68). */
static void finish_path(const otrk_opts *o, int code, const ocplx *fpm, const ocplx *gamma, ocplx *Yf, ocplx *fpt,
                        otrk_stats *st) {
    ocplx u_dehom[ONY];
    memset(u_dehom, 0, sizeof(u_dehom));
    u_dehom[ONV] = C(1., 0);
    oracle_switch_patch(u_dehom, Yf, Yf);
    if (code == 0 && o->polish_its > 0) polish_root(o, u_dehom, fpm, gamma, Yf, st);
    ocplx cogn[2 * ONV];
    oracle_cognate_solns(Yf, fpm, cogn);
    memcpy(fpt, cogn, sizeof(ocplx) * ONV);
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) return nthreads;
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

int oracle_homotopy(const otrk_opts *o, const ocplx *fpm, ocplx gamma, uint64_t seed,
                    uint64_t first_path_id, int32_t n, ocplx *fpt, int32_t *err, otrk_stats *stats,
                    ocplx *yfinal, int nthreads) {
    int nt = set_threads(nthreads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int32_t i = 0; i < n; ++i) {
        ocplx spt[ONV], spm[ONP], Yf[ONY];
        otrk_stats st;
        uint64_t pid = first_path_id + (uint64_t)i;
        oracle_generate_start(o, seed, pid, spt, spm);
        err[i] = oracle_track_path(o, spt, spm, fpm, &gamma, seed, pid, 0, Yf, &st);
        if (yfinal) memcpy(&yfinal[(int64_t)i * ONY], Yf, sizeof(Yf));
        finish_path(o, err[i], fpm, &gamma, Yf, &fpt[(int64_t)i * ONV], &st);
        if (stats) stats[i] = st;
    }
    return 0;
}

int oracle_homotopy_explicit(const otrk_opts *o, const ocplx *fpm, ocplx gamma, const ocplx *spt,
                             const ocplx *spm, uint64_t seed, uint64_t first_path_id, int32_t n,
                             ocplx *fpt, int32_t *err, otrk_stats *stats, ocplx *yfinal,
                             int nthreads) {
    int nt = set_threads(nthreads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int32_t i = 0; i < n; ++i) {
        ocplx Yf[ONY];
        otrk_stats st;
        uint64_t pid = first_path_id + (uint64_t)i;
        err[i] = oracle_track_path(o, &spt[(int64_t)i * ONV], &spm[(int64_t)i * ONP], fpm, &gamma,
                                   seed, pid, 0, Yf, &st);
        if (yfinal) memcpy(&yfinal[(int64_t)i * ONY], Yf, sizeof(Yf));
        finish_path(o, err[i], fpm, &gamma, Yf, &fpt[(int64_t)i * ONV], &st);
        if (stats) stats[i] = st;
    }
    return 0;
}

int oracle_homotopy_explicit_ids(const otrk_opts *o, const ocplx *fpm, ocplx gamma, const ocplx *spt,
                                 const ocplx *spm, uint64_t seed, const uint64_t *path_ids, int32_t n,
                                 ocplx *fpt, int32_t *err, int nthreads) {
    int nt = set_threads(nthreads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int32_t i = 0; i < n; ++i) {
        ocplx Yf[ONY];
        otrk_stats st;
        err[i] = oracle_track_path(o, &spt[(int64_t)i * ONV], &spm[(int64_t)i * ONP], fpm, &gamma, seed,
                                   path_ids[i], 0, Yf, &st);
        finish_path(o, err[i], fpm, &gamma, Yf, &fpt[(int64_t)i * ONV], &st);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* L7: host post-processing (frgsupport.cu)                                                    */
/* ------------------------------------------------------------------------------------------ */

/* LambertW principal branch by Halley's method (frgsupport.cu:220-272). */
static double zexpz(double z) { return z * exp(z); }
static double zexpz_d(double z) { return exp(z) + z * exp(z); }
static double zexpz_dd(double z) { return 2. * exp(z) + z * exp(z); }
static double lw_init(double z) {
    const double e = 2.71828182845904523536;
    double ip = log(z) - log(log(z));
    double p = sqrt(2. * (e * z + 1.));
    if (fabs(z - (-exp(-1.))) <= 1.) ip = -1. + p - 1. / 3. * pow(p, 2) + 11. / 72. * pow(p, 3);
    if (fabs(z - .5) <= .5) ip = (0.35173371 * (0.1237166 + 7.061302897 * z)) / (2. + 0.827184 * (1. + 2. * z));
    return ip;
}
double oracle_lambertw(double z) {
    if (z == 0.) return 0.;
    if (z == -exp(-1.)) return -1.;
    if (z == exp(1.)) return 1.;
    double w = lw_init(z), wprev;
    unsigned iter = 0;
    do {
        wprev = w;
        const double dw = 2. * ((zexpz(w) - z) * zexpz_d(w)) / (2. * pow(zexpz_d(w), 2) - (zexpz(w) - z) * zexpz_dd(w));
        w = MUT(13) ? w + dw : w - dw;
        iter++;
    } while ((fabs(w - wprev) > 1.E-30) && iter < 30);
    return w;
}

/* SortingDeletingMerging (frgsupport.cu:10-215), restated without the list/deque mechanics:
 *  1. order the new roots by Re(z_0) (ascending; the reference sorts on the device at
 *     frg.cu:432-435; here a stable sort with NaN keys last);
 *  2. merge with the collection, old roots first on equal keys (frgsupport.cu:34-44);
 *  3. a new root with code 0 becomes a duplicate (code 6) if, walking left and right from its
 *     merged rank while |dRe z_0| <= tol, it meets an old root, or a new root with smaller
 *     index that still has code 0, whose 42 reals are all within tol (:69-171);
 *  4. the merged, sorted collection keeps old roots and new code-0 roots (:191-211). */
typedef struct { double key; int64_t idx; } okey;
static int okey_cmp(const void *a, const void *b) {
    const okey *x = (const okey *)a, *y = (const okey *)b;
    int xn = isnan(x->key), yn = isnan(y->key);
    if (xn != yn) return xn - yn;
    if (!xn && x->key < y->key) return -1;
    if (!xn && x->key > y->key) return 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}
static int same_root(const ocplx *a, const ocplx *b, double tol) {
    for (int j = 0; j < (MUT(14) ? ONV - 1 : ONV); ++j) {
        if (fabs(a[j].re - b[j].re) > tol) return 0;
        if (fabs(a[j].im - b[j].im) > tol) return 0;
    }
    return 1;
}
int64_t oracle_sort_dedupe_merge(const ocplx *coll, int64_t m, const ocplx *fpt, int32_t *err,
                                 int64_t n, double tol, ocplx *coll_out) {
    okey *ord = (okey *)malloc(sizeof(okey) * (size_t)(n > 0 ? n : 1));
    int64_t *NtoT = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + n + 1));
    int64_t *TtoN = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + n + 1));
    for (int64_t i = 0; i < n; ++i) { ord[i].key = fpt[i * ONV].re; ord[i].idx = i; }
    qsort(ord, (size_t)n, sizeof(okey), okey_cmp);
    /* merge: global trial numbers 0..m-1 are old roots, m+i are new path i */
    int64_t io = 0, in = 0, r = 0;
    while (io < m || in < n) {
        if (io < m && (in >= n || !(fpt[ord[in].idx * ONV].re < coll[io * ONV].re))) NtoT[r++] = io++;
        else NtoT[r++] = m + ord[in++].idx;
    }
    for (int64_t i = 0; i < m + n; ++i) TtoN[NtoT[i]] = i;
    for (int64_t t = 0; t < n; ++t) {
        if (err[t] != 0) continue;
        const ocplx *ft = &fpt[t * ONV];
        int64_t rank = TtoN[m + t];
        for (int dir = -1; dir <= 1; dir += 2) {
            for (int64_t i = rank + dir; i >= 0 && i < m + n; i += dir) {
                int64_t ug = NtoT[i], ul = ug - m;
                const ocplx *fu = ul < 0 ? &coll[ug * ONV] : &fpt[ul * ONV];
                if (fabs(ft[0].re - fu[0].re) > tol) break;
                if (ul < 0) {
                    if (same_root(ft, fu, tol)) err[t] = 6;
                } else if (ul < t && err[ul] == 0) {
                    if (same_root(ft, fu, tol)) err[t] = 6;
                }
            }
        }
    }
    int64_t k = 0;
    for (int64_t i = 0; i < m + n; ++i) {
        int64_t g = NtoT[i];
        if (g < m) memcpy(&coll_out[k++ * ONV], &coll[g * ONV], sizeof(ocplx) * ONV);
        else if (err[g - m] == 0) memcpy(&coll_out[k++ * ONV], &fpt[(g - m) * ONV], sizeof(ocplx) * ONV);
    }
    free(ord); free(NtoT); free(TtoN);
    return k;
}
