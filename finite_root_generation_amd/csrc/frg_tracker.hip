// frg_tracker.hip -- the MI355X (gfx950) homotopy path tracker.
//
// Replaces the reference's Homotopy kernel (frg.cu:35-69) and everything it inlines:
// GenerateStart (frgkinematics.cu:13-136), TrackPath (htpytrack.cu:33-220), DormandPrince65
// (rungekutta.cu:3-124), IVP/Newton/LinearSolve* (htpysupport.cu:146-494), FSYS/JVAR/JPAR
// (htpykinematics.cu) and CognateSolns (frgkinematics.cu:227-285).
//
// Mapping.  One path per 8-lane group, eight paths per wave64.  Lane r < 7 owns pose j = r+1:
// it evaluates equations I_j, II_j, III_j and their Jacobian rows.  Lane 7 owns the two patch
// rows (projective patch u and velocity/time row w).  A 23-vector Y = [z(21), Z0, t] is held
// as three complex slots per lane: lane r < 7 holds z[2r], z[2r+1], U_{r+1} (= z[14+r]); lane 7
// holds Z0, t and a zero.
//
// Linear solves.  Every tangent (IVP) and Newton step solves the 23x23 system
//     [ JH (rows 0..21, row-scaled) ; w ] x = b
// (htpysupport.cu:250-283, :431-477).  U_j appears only in the three rows of pose j and in the
// two patch rows, so the solve is done in three stages instead of one dense LU:
//   1. in-lane: pivot among the three rows of pose j on |dF/dU_j|, eliminate U_j, leaving two
//      rows over the 16 shared columns (14 linkage, Z0, t) and one pivot row;
//   2. eliminate U_1..U_7 from the two patch rows (columns split over the 8 lanes via LDS);
//   3. Gauss-Jordan with partial pivoting on the remaining 16x16 system, two rows per lane,
//      pivot search by DPP arg-max, pivot row broadcast through LDS; then back-substitute U_j.
// The Newton system (LinearSolveNewton, htpysupport.cu:302-400) is the same solve with w = e_t
// and right-hand side dt = tgoal - t, which pins x_t exactly as the reference does.
//
// Control flow.  Each group runs TrackPath as a state machine whose every step is one solve, so
// all groups of a wave execute the expensive part together whatever phase their path is in;
// only the cheap bookkeeping diverges.  A persistent variant refills a group from a device work
// queue when its path ends.
#include "frg_device.h"
#include "frg_system.h"
#include "frg_kernels.h"
#include "frg_solve16.h"

namespace frg {

// ------------------------------------------------------------------------------------------
// RK6(5)8M tableau (rungekutta.cu:43-122).  kA[i][m]: stage i point = Yc + s * sum_m kA[i][m] k_m.
// ------------------------------------------------------------------------------------------
__constant__ double kA[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {1. / 10., 0, 0, 0, 0, 0, 0},
    {-2. / 81., 20. / 81., 0, 0, 0, 0, 0},
    {615. / 1372., -270. / 343., 1053. / 1372., 0, 0, 0, 0},
    {3243. / 5500., -54. / 55., 50949. / 71500., 4998. / 17875., 0, 0, 0},
    {-26492. / 37125., 72. / 55., 2808. / 23375., -24206. / 37125., 338. / 459., 0, 0},
    {5561. / 2376., -35. / 11., -24117. / 31603., 899983. / 200772., -5225. / 1836., 3925. / 4056., 0},
    {465467. / 266112., -2945. / 1232., -5610201. / 14158144., 10513573. / 3212352., -424325. / 205632.,
     376225. / 454272., 0}};
// error (rungekutta.cu:99-105) and propagated solution (:115-122) weights; k2 enters neither
__constant__ double kE[8] = {-13. / 2400., 0., 19683. / 618800., -2401. / 31200., 65. / 816., -15. / 416.,
                              -521. / 5600., 1. / 10.};
__constant__ double kB[8] = {61. / 864., 0., 98415. / 321776., 16807. / 146016., 1375. / 7344., 1375. / 5408.,
                              -37. / 1120., 1. / 10.};

// ------------------------------------------------------------------------------------------
// Per-group LDS scratch
// ------------------------------------------------------------------------------------------
template <typename R>
struct GroupLds {
    cx<R> phat[7][17];  // normalised pivot row of each pose (16 shared columns + rhs)
    cx<R> uw[14];       // U-entries of the (scaled) u row [0..6] and of the w row [7..13]
    cx<R> a[32];        // eval point (16) / reduced patch rows (32) / pivot row (17) / solution (16)
};


// Y-vector index held in slot k of lane `role` (-1: the unused third slot of lane 7).
__device__ __forceinline__ int slot_index(int role, int k) {
    if (role < 7) return k == 2 ? 14 + role : 2 * role + k;
    return k == 2 ? -1 : 21 + k;
}

// Row scaling by 1/max(|re|,|im|) over the row (htpysupport.cu:260-277, :446-471); the Newton
// right-hand side -F is scaled by the same factor.  kOwnF: F comes with the row (FP64);
// otherwise it is the FP64 residual fext.
template <typename R, bool kOwnF>
__device__ __forceinline__ void scale_row(const Row<R> &r, bool newton, cx<double> fext, cx<R> g[16], cx<R> &au,
                                          cx<R> &rhs) {
    R m = fmax(fabs(r.au.re), fabs(r.au.im));
#pragma unroll
    for (int k = 0; k < 16; ++k) m = fmax(m, fmax(fabs(r.g[k].re), fabs(r.g[k].im)));
    const R s = recip(m);
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = s * r.g[k];
    au = s * r.au;
    const cx<double> f = sel(kOwnF, cvt<double>(r.f), fext);
    rhs = sel(newton, cvt<R>(mk(-f.re * double(s), -f.im * double(s))), mk(R(0), R(0)));
}

// ------------------------------------------------------------------------------------------
// The group solve.  Inputs are this lane's slots of the evaluation point Ye, the patch u and the
// last row w; `newton` selects the Newton right-hand side (-H scaled; w rhs = dt) over the
// tangent one (zeros; w rhs = 1).  fnewt (FP64 residual values of this lane's three
// equations) is used when `newton`.  Output: this lane's slots of x.
// ------------------------------------------------------------------------------------------
// kExtRes: the Newton residual comes from fnewt (always for FP32 rows, whose F is not accurate
// enough; for FP64 rows when the caller evaluates it more accurately) instead of with the rows.
template <typename R, bool kExtRes = (sizeof(R) != sizeof(double))>
__device__ __forceinline__ void group_solve(int role, GroupLds<R> *L, const cx<double> Ye[3], const cx<double> uu[3],
                                            const cx<double> ww[3], bool newton, double w_rhs,
                                            const cx<double> *par, cx<double> gamma,
                                            const cx<double> fnewt[3], cx<double> x[3]) {
    constexpr bool kOwnF = !kExtRes;
    const cx<R> zero = mk(R(0), R(0));
    const int pr = role < 7 ? role : 6;  // LDS row index used by this lane for pose data

    // -------- patch-row scale and residual (group reductions, FP64) --------
    double um = 0.0;
    cx<double> uy = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        um = fmax(um, fmax(fabs(uu[k].re), fabs(uu[k].im)));
        uy = fma(uy, uu[k], Ye[k]);
    }
    um = grp_max(um);
    uy = grp_sum(uy);
    const double usc = 1.0 / um;
    const cx<double> bu = sel(newton, mk(-(uy.re - 1.0) * usc, -uy.im * usc), mk(0.0, 0.0));

    // -------- 1. broadcast the shared part of the evaluation point --------
    L->a[2 * role] = cvt<R>(Ye[0]);
    L->a[2 * role + 1] = cvt<R>(Ye[1]);
    lds_sync();
    PoseEval<R> e;
#pragma unroll
    for (int k = 0; k < 14; ++k) e.L[k] = L->a[k];
    e.Z0 = L->a[14];
    e.t = L->a[15].re;
    lds_sync();
    e.U = sel(role < 7, cvt<R>(Ye[2]), mk(R(1), R(0)));
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        e.s[p] = cvt<R>(par[p]);
        e.d[p] = cvt<R>(par[6 + p]);
    }
    e.gamma = cvt<R>(gamma);

    // -------- 2. pose rows one at a time: row scaling, pairwise pivoting on |dF/dU_j| --------
    // Rows I_j, II_j, III_j are produced in turn.  p holds the pivot candidate (largest |dF/dU_j|
    // so far, the first one on ties, as LinearSolve's search); each other row is reduced by it,
    // with multipliers of modulus <= 1.  The final p is the row of largest |dF/dU_j|.
    PoseCommon<R> pc;
    pose_common(e, pc);
    cx<R> pg[16], pau, prhs;
    cx<R> ra[17], rb[17];
    {
        Row<R> r;
        pose_row<0>(e, pc, r);
        scale_row<R, kOwnF>(r, newton, fnewt[0], pg, pau, prhs);
    }
#pragma unroll
    for (int eq = 1; eq < 3; ++eq) {
        Row<R> r;
        if (eq == 1) pose_row<1>(e, pc, r);
        else pose_row<2>(e, pc, r);
        cx<R> g[16], au, rhs;
        scale_row<R, kOwnF>(r, newton, fnewt[eq], g, au, rhs);
        const bool swap = abs2(au) > abs2(pau);
        // (p, r) <- swap ? (r, p) : (p, r), then r <- r - (r_U / p_U) p
        cx<R> red[17];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const cx<R> pk = sel(swap, g[k], pg[k]);
            red[k] = sel(swap, pg[k], g[k]);
            pg[k] = pk;
        }
        red[16] = sel(swap, prhs, rhs);
        prhs = sel(swap, rhs, prhs);
        const cx<R> rU = sel(swap, pau, au);
        pau = sel(swap, au, pau);
        const cx<R> m = rU * inv(pau);
#pragma unroll
        for (int k = 0; k < 16; ++k) red[k] = fms(red[k], m, pg[k]);
        red[16] = fms(red[16], m, prhs);
#pragma unroll
        for (int k = 0; k < 17; ++k) {
            if (eq == 1) ra[k] = red[k];
            else rb[k] = red[k];
        }
    }
    const cx<R> ipau = inv(pau);
    // normalised pivot row -> LDS; U-entries of u (scaled) and w -> LDS
    const cx<R> uU = cvt<R>(usc * uu[2]);
    const cx<R> wU = cvt<R>(ww[2]);
    const cx<R> rho = prhs * ipau;
    if (role < 7) {
#pragma unroll
        for (int k = 0; k < 16; ++k) L->phat[pr][k] = pg[k] * ipau;
        L->phat[pr][16] = rho;
        L->uw[pr] = uU;
        L->uw[7 + pr] = wU;
    }
    lds_sync();

    // -------- 3. eliminate U_1..U_7 from the patch rows; lane r builds columns 2r, 2r+1 --------
    {
        cx<R> cu0 = cvt<R>(usc * uu[0]), cu1 = cvt<R>(usc * uu[1]);
        cx<R> cw0 = cvt<R>(ww[0]), cw1 = cvt<R>(ww[1]);
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const cx<R> a0 = L->phat[j][2 * role], a1 = L->phat[j][2 * role + 1];
            const cx<R> mu = L->uw[j], mw = L->uw[7 + j];
            cu0 = fms(cu0, mu, a0);
            cu1 = fms(cu1, mu, a1);
            cw0 = fms(cw0, mw, a0);
            cw1 = fms(cw1, mw, a1);
        }
        lds_sync();
        L->a[2 * role] = cu0;
        L->a[2 * role + 1] = cu1;
        L->a[16 + 2 * role] = cw0;
        L->a[16 + 2 * role + 1] = cw1;
    }
    // right-hand sides of the reduced patch rows: b - sum_j (U-entry_j * rho_j)
    const cx<R> su = grp_sum(sel(role < 7, uU * rho, zero));
    const cx<R> sw = grp_sum(sel(role < 7, wU * rho, zero));
    lds_sync();
    if (role == 7) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            ra[k] = L->a[k];
            rb[k] = L->a[16 + k];
        }
        ra[16] = cvt<R>(bu) - su;
        rb[16] = mk(R(w_rhs), R(0)) - sw;
    }
    lds_sync();

    // -------- 4. Gauss-Jordan on the 16 x 16 system (rows 2*role, 2*role+1) --------
    // Pivot choice per column: the largest |a|^2 among the unused rows (pivot_key packs the
    // magnitude and the row id into one word, so the group search is three integer maxima).
    // The pivot lane broadcasts 1/pivot and the rest of its row through LDS.
    bool usedA = false, usedB = false;
    int pcA = 0, pcB = 0;
    cx<R> invA = zero, invB = zero;
    const int idA = 2 * role, idB = 2 * role + 1;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const uint32_t key = max(pivot_key(abs2(ra[c]), usedA, idA), pivot_key(abs2(rb[c]), usedB, idB));
        const int gid = 15 - (int)(grp_max_u32(key) & 15u);
        const bool pA = gid == idA, pB = gid == idB;
        const cx<R> ipiv = inv(sel(pB, rb[c], ra[c]));
        if (pA) {
            usedA = true;
            pcA = c;
            invA = ipiv;
        }
        if (pB) {
            usedB = true;
            pcB = c;
            invB = ipiv;
        }
        // one write branch for either row of the pivot lane: every LDS instruction of the wave
        // occupies the CU's LDS pipe, so two exec-masked branches of (16 - c) writes each cost
        // twice the issue time of one
        if (pA || pB) {
            L->a[c] = ipiv;
#pragma unroll
            for (int k = 0; k < 17; ++k)
                if (k > c) L->a[k] = sel(pB, rb[k], ra[k]);
        }
        lds_order();
        const cx<R> ip = L->a[c];
        const cx<R> la = sel(pA, zero, ra[c] * ip);
        const cx<R> lb = sel(pB, zero, rb[c] * ip);
#pragma unroll
        for (int k = 0; k < 17; ++k) {
            if (k > c) {
                const cx<R> pk = L->a[k];
                ra[k] = fms(ra[k], la, pk);
                rb[k] = fms(rb[k], lb, pk);
            }
        }
        // the next column's pivot lane overwrites L->a only after these reads: LDS requests of
        // a wave complete in issue order, and the compiler keeps stores after loads of the same
        // addresses
        lds_order();
    }
    lds_sync();
    // -------- 5. solution of the shared unknowns --------
    L->a[pcA] = ra[16] * invA;
    L->a[pcB] = rb[16] * invB;
    lds_sync();
    x[0] = cvt<double>(L->a[2 * role]);
    x[1] = cvt<double>(L->a[2 * role + 1]);
    // -------- 6. back-substitute U_j = rho - phat . x --------
    {
        cx<R> acc = L->phat[pr][16];
#pragma unroll
        for (int k = 0; k < 16; ++k) acc = fms(acc, L->phat[pr][k], L->a[k]);
        x[2] = sel(role < 7, cvt<double>(acc), mk(0.0, 0.0));
    }
    lds_sync();
}

// ------------------------------------------------------------------------------------------
// Tracker state
// ------------------------------------------------------------------------------------------
enum Phase : int { PH_INIT = 0, PH_RK = 1, PH_NEWTON = 2, PH_SHARP = 3 };

// Per-path values that are the same in every lane of a group and rarely read.
struct alignas(16) GroupScalars {
    cx<double> gamma;
    uint64_t pid, seed;  // RNG stream of the path
    int64_t path;        // global path index (output row)
    int inst;            // task instance
    uint32_t draw;       // next patch draw
    int rec;             // step log: row of the last proposal's record, or -1
    // statistics (Newton calls = ministep - st_rk)
    int st_rk, st_newton_its, st_redraws, st_fp32, st_tan;
    // TrackPath state read once per RK attempt or Newton call
    double err, scale;  // last RK error estimate and its step factor
    int code, step, ministep, streak, attempt, rej;
};

// Per-group LDS of the 8-lane layout: the solve scratch plus state that is read once per solve
// and so lives in LDS rather than in VGPRs (the kernel's occupancy is register-bound).
struct GroupSlot {
    GroupLds<double> lds;   // also used as GroupLds<float> by FP32 solves
    cx<double> par[8][12];  // per lane: start values s6[0..5] and target - start d6[0..5] of its pose
    cx<double> vc[8][3];    // per lane: its slots of the current velocity Vc
    cx<double> u[8][3];     // per lane: its slots of the projective patch u
    GroupScalars gs;
    // The groups of a wave read the same field of their own slots together (a broadcast within
    // each group): with a slot size that is an odd multiple of 16 bytes those 16-byte reads fall
    // in disjoint LDS banks (8 groups x 16 B over 32 or 64 four-byte banks).  5,072 bytes: 32
    // slots still fit the 160 KB of a CU.
    unsigned char bank_pad[16];
};
static_assert(sizeof(GroupSlot) % 16 == 0 && offsetof(GroupLds<double>, a) % 16 == 0,
              "the step model reads float4 vectors from GroupLds::a");
static_assert((sizeof(GroupSlot) / 16) % 2 == 1, "slot stride: an odd number of 16-byte bank rows");
static_assert(32 * sizeof(GroupSlot) <= 160 * 1024, "32 groups (256 threads) per CU");
static_assert(sizeof(GroupLds<double>::a) >= 128 * sizeof(float), "features (8) and hidden layer (64 at +64)");

// Per-group LDS of the 16-lane layout (4,752 bytes: 32 paths on a CU at two waves per SIMD).
struct GroupSlot16 {
    GroupLds16 lds;
    cx<double> par[60];    // the 30 parameters as (start value, target - start) pairs
    cx<double> vc[16][2];  // per lane: its slots of the current velocity Vc
    cx<double> u[16][2];   // per lane: its slots of the projective patch u
    GroupScalars gs;
    unsigned char bank_pad[16];  // odd number of 16-byte bank rows, as GroupSlot
};
static_assert(sizeof(GroupSlot16) % 16 == 0 && offsetof(GroupLds16, a) % 16 == 0,
              "the step model reads float4 vectors from GroupLds16::a");
static_assert((sizeof(GroupSlot16) / 16) % 2 == 1, "slot stride: an odd number of 16-byte bank rows");
static_assert(32 * sizeof(GroupSlot16) <= 160 * 1024, "32 groups per CU at two waves per SIMD");
static_assert(sizeof(GroupLds16::a) >= 128 * sizeof(float), "features (8) and hidden layer (64 at +64)");

// Lane layouts.  kW lanes track one path; each lane holds kS slots of every 23-vector
// Y = [z(21), Z0, t]; t sits in slot kTSlot of lane kTLane, Z0 in slot kZSlot of lane kZLane.
struct Lanes8 {
    static constexpr int kW = 8, kS = 3, kTLane = 7, kTSlot = 1, kZLane = 7, kZSlot = 0;
    using Slot = GroupSlot;
    // lane r < 7: z[2r], z[2r+1], U_{r+1}; lane 7: Z0, t, -
    __device__ static int slot_index(int role, int k) {
        if (role < 7) return k == 2 ? 14 + role : 2 * role + k;
        return k == 2 ? -1 : 21 + k;
    }
};
struct Lanes16 {
    static constexpr int kW = 16, kS = 2, kTLane = 15, kTSlot = 0, kZLane = 14, kZSlot = 0;
    using Slot = GroupSlot16;
    // lane 2p+h < 14: z[7h+p] and (h = 0) U_{p+1}; lane 14: Z0; lane 15: t
    __device__ static int slot_index(int role, int k) {
        if (role < 14) return k == 0 ? 7 * (role & 1) + (role >> 1) : ((role & 1) ? -1 : 14 + (role >> 1));
        return k == 0 ? 7 + role : -1;
    }
};

template <class G>
struct PathState {
    // vectors (this lane's slots)
    cx<double> Yc[G::kS], K[8][G::kS];  // K[0] doubles as YN, K[1] as Yn after an RK attempt
    cx<double> *u;                       // this lane's row of the group's u
    cx<double> *par;                     // 8 lanes: this lane's pose parameters; 16: the group's 30
    cx<double> *vc;                      // this lane's row of the group's vc
    GroupScalars *gs;
    // scalars (identical in all lanes of the group)
    double s, residual;
    union {
        double tgoal;      // Newton target t
        double s_classic;  // classical proposal of a step waiting for the model (the fallback)
    };
    int phase, stage, its;
    bool tstep, done;
    bool pending;  // learned step control: an accepted step waits for the model's proposal
};

// Patch draw (htpytrack.cu:55-58,147-150): Y index i uses draws base + 2i, base + 2i + 1.
template <class G>
__device__ __forceinline__ void draw_vec(int role, const PathState<G> &P, uint64_t base, double lo, double hi,
                                         cx<double> out[G::kS]) {
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        const int yi = G::slot_index(role, k);
        double a, b;
        uniform_pair(base / 2 + (uint64_t)(yi < 0 ? 0 : yi), P.gs->pid, P.gs->seed, a, b);
        out[k] = sel(yi < 0, mk(0.0, 0.0), mk((hi - lo) * a + lo, (hi - lo) * b + lo));
    }
}

template <class G>
__device__ __forceinline__ bool is_t_slot(int role, int k) { return role == G::kTLane && k == G::kTSlot; }

// SwitchPatch (htpysupport.cu:9-34): Y' = (1 - u_t t)/(sum_{i<22} u_i Y_i) * Y, t unchanged.
template <class G>
__device__ __forceinline__ void switch_patch(int role, const cx<double> u[G::kS], cx<double> Y[G::kS]) {
    cx<double> den = mk(0.0, 0.0), num = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        const bool is_t = is_t_slot<G>(role, k);
        den = sel(is_t, den, fma(den, u[k], Y[k]));
        num = sel(is_t, u[k] * Y[k], num);
    }
    den = grp_sum<G::kW>(den);
    num = grp_sum<G::kW>(num);
    const cx<double> sc = cdiv(mk(1.0 - num.re, -num.im), den);
#pragma unroll
    for (int k = 0; k < G::kS; ++k) Y[k] = sel(is_t_slot<G>(role, k), Y[k], sc * Y[k]);
}

template <class G>
__device__ __forceinline__ double grp_norm(const cx<double> v[G::kS]) {
    double a = abs2(v[0]);
#pragma unroll
    for (int k = 1; k < G::kS; ++k) a = a + abs2(v[k]);
    return sqrt(grp_sum<G::kW>(a));
}

// Tangent normalisation (htpysupport.cu:285-296).  Arclength: x * e^{-i arg x_t} / |x|, written
// as x * conj(x_t)/(|x_t| |x|); t-step: x / x_t.  Then Im x_t = 0.
template <class G>
__device__ __forceinline__ void normalize_tangent(int role, cx<double> x[G::kS], bool tstep) {
    const double nrm = grp_norm<G>(x);
    const cx<double> xt = grp_bcast<G::kW>(x[G::kTSlot], role, G::kTLane);
    cx<double> sc;
    if (!tstep) {
        const double a = sqrt(abs2(xt));
        sc = sel(a > 0.0, mk(xt.re / (a * nrm), -xt.im / (a * nrm)), mk(1.0 / nrm, 0.0));
    } else {
        sc = inv(xt);
    }
#pragma unroll
    for (int k = 0; k < G::kS; ++k) x[k] = sc * x[k];
    if (role == G::kTLane) x[G::kTSlot].im = 0.0;
}
// t (real part of Y[22]), delivered to the whole group
template <class G>
__device__ __forceinline__ double grp_t(int role, const cx<double> Y[G::kS]) {
    return grp_bcast<G::kW>(Y[G::kTSlot].re, role, G::kTLane);
}

// ------------------------------------------------------------------------------------------
// Start system: GenerateStart (frgkinematics.cu:13-136) for pose j (the lane's pose; a lane
// without a pose repeats pose 7).
// ------------------------------------------------------------------------------------------
template <class G>
__device__ __forceinline__ cx<double> start_value(const TrackOpts &o, uint64_t blk, const PathState<G> &P) {
    double a, b;
    uniform_pair(blk, P.gs->pid, P.gs->seed, a, b);
    if (o.start_mode == 1) {
        const double r = sqrt(-log(a)), th = 6.283185307179586 * b;
        double sn, cs;
        sincos(th, &sn, &cs);
        return mk(r * cs, r * sn);
    }
    return mk((o.start_hi - o.start_lo) * a + o.start_lo, (o.start_hi - o.start_lo) * b + o.start_lo);
}

// The start point's linkage z[0..13] and pose j's rotation U and parameters
// (P0, P_j, Pc0, Pc_j, Q_j, Qc_j).
struct StartPose {
    cx<double> z[14], U, q[6];
};

template <class G>
__device__ __forceinline__ void start_pose(const TrackOpts &o, const PathState<G> &P, int j, StartPose &sp) {
    const cx<double> zero = mk(0.0, 0.0);
    cx<double> *z = sp.z;
#pragma unroll
    for (int i = 0; i < 14; ++i) z[i] = start_value(o, (uint64_t)i, P);
    const cx<double> P0 = start_value(o, 14, P), Pc0 = start_value(o, 15, P);
    const cx<double> R = start_value(o, 15 + (uint64_t)j, P), iR = inv(R);
    double pmA, pmB;
    uniform_pair(22 + (uint64_t)j, P.gs->pid, P.gs->seed, pmA, pmB);
    const cx<double> A = z[0], B = z[1], C = z[2], D = z[3], F = z[4], G_ = z[5], H = z[6];
    const cx<double> Ac = z[7], Bc = z[8], Cc = z[9], Dc = z[10], Fc = z[11], Gc = z[12], Hc = z[13];
    cx<double> Aq = (G_ - D) * (iR * (Hc - Gc) - (Hc - Fc));
    cx<double> Acq = (Gc - Dc) * (R * (H - G_) - (H - F));
    cx<double> Bq = ((zero - R * ((H - G_) * (Hc - Fc))) - iR * ((Hc - Gc) * (H - F))) +
                    ((G_ - D) * (Gc - Dc) + ((H - G_) * (Hc - Gc) + ((H - F) * (Hc - Fc) - (F - D) * (Fc - Dc))));
    double pm = pmA < 0.5 ? -1.0 : 1.0;
    const cx<double> TUc = inv(mk(2.0, 0.0) * Aq) *
                           ((zero - Bq) + mk(pm, 0.0) * csqrt_ref(Bq * Bq - mk(4.0, 0.0) * (Aq * Acq)));
    const cx<double> SUc = inv(F - D) * (TUc * (G_ - D) + (R * (H - G_) - (H - F)));
    const cx<double> iSUc = inv(SUc);
    const cx<double> E = (R * (H - C) - SUc * (F - B)) - (H - F);
    const cx<double> Ec = (iR * (Hc - Cc) - iSUc * (Fc - Bc)) - (Hc - Fc);
    Aq = (Ac - Bc) * E;
    Acq = (A - B) * Ec;
    Bq = E * Ec + ((A - B) * (Ac - Bc) - (C - A) * (Cc - Ac));
    pm = pmB < 0.5 ? -1.0 : 1.0;
    const cx<double> U = inv(mk(2.0, 0.0) * Aq) *
                         ((zero - Bq) + mk(pm, 0.0) * csqrt_ref(Bq * Bq - mk(4.0, 0.0) * (Aq * Acq)));
    const cx<double> QUc = mk(-1.0, 0.0) * inv(C - A) * (inv(U) * (A - B) + E);
    const cx<double> Q = QUc * U, Qc = inv(Q);
    const cx<double> S = SUc * U;
    sp.U = U;
    sp.q[0] = P0;
    sp.q[1] = B + (S * (F - B) + U * (P0 - F));
    sp.q[2] = Pc0;
    sp.q[3] = Bc + (inv(S) * (Fc - Bc) + inv(U) * (Pc0 - Fc));
    sp.q[4] = Q;
    sp.q[5] = Qc;
}

// Pose of a lane: 8 lanes: lane r < 7 -> pose r+1; 16 lanes: lane 2p+h < 14 -> pose p+1; the
// patch lanes repeat pose 7.
template <class G>
__device__ __forceinline__ int lane_pose(int role) {
    if (G::kW == 8) return role < 7 ? role + 1 : 7;
    return role < 14 ? (role >> 1) + 1 : 7;
}

// Start values of this lane's slots and parameters: GenerateStart (frgkinematics.cu:13-136), or
// the caller's explicit start system (spt, spm rows).  8 lanes: the lane's six pose parameters
// go to par[0..5]; 16 lanes: the group's 30 parameters to par[2i] (two per lane).
template <class G>
__device__ __forceinline__ void start_values(int role, const Job &J, PathState<G> &P, int64_t path) {
    const int j = lane_pose<G>(role);
    const int pidx[6] = {0, j, 8, 8 + j, 15 + j, 22 + j};  // P0, P_j, Pc0, Pc_j, Q_j, Qc_j
    cx<double> y[G::kS];
    if (J.spt) {
        const cx<double> *zs = J.spt + path * 21;
        const cx<double> *qs = J.spm + path * 30;
#pragma unroll
        for (int k = 0; k < G::kS; ++k) {
            const int yi = G::slot_index(role, k);
            y[k] = yi < 0 ? mk(0.0, 0.0) : (yi < 21 ? zs[yi] : mk(yi == 21 ? 1.0 : 0.0, 0.0));
        }
        if (G::kW == 8) {
#pragma unroll
            for (int p = 0; p < 6; ++p) P.par[p] = qs[pidx[p]];
        } else {
            P.par[2 * role] = qs[role];
            if (role + 16 < 30) P.par[2 * (role + 16)] = qs[role + 16];
        }
    } else {
        StartPose sp;
        start_pose(J.o, P, j, sp);
#pragma unroll
        for (int k = 0; k < G::kS; ++k) {
            const int yi = G::slot_index(role, k);
            cx<double> v = mk(yi == 21 ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int i = 0; i < 14; ++i) v = sel(yi == i, sp.z[i], v);
            y[k] = sel(yi >= 14 && yi < 21, sp.U, v);
        }
        if (G::kW == 8) {
#pragma unroll
            for (int p = 0; p < 6; ++p) P.par[p] = sp.q[p];
        } else if (role < 14) {
            const int h = role & 1;
            // the even lane of pose j stores P_j and Q_j, the odd one Pc_j and Qc_j; lanes 0 and
            // 1 also P0 and Pc0
            P.par[2 * pidx[1 + 2 * h]] = sp.q[1 + 2 * h];
            P.par[2 * pidx[4 + h]] = sp.q[4 + h];
            if (role < 2) P.par[2 * pidx[2 * h]] = sp.q[2 * h];
        }
    }
#pragma unroll
    for (int k = 0; k < G::kS; ++k) P.Yc[k] = y[k];
}

// Task instance of a global path index: the last instance whose first path is <= path.
__device__ __forceinline__ int instance_of(const Job &J, int64_t path) {
    if (J.n_inst <= 1) return 0;
    int lo = 0, hi = J.n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J.inst_first[mid] <= path) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------
// Path set-up (Homotopy kernel start + TrackPath initialisation, htpytrack.cu:45-77)
// ------------------------------------------------------------------------------------------
template <class G>
__device__ __forceinline__ void init_path(int role, const Job &J, PathState<G> &P, int64_t path) {
    const TrackOpts &o = J.o;
    P.gs->path = path;
    const int inst = instance_of(J, path);
    P.gs->inst = inst;
    const int64_t local = path - (J.n_inst > 1 ? J.inst_first[inst] : 0);
    P.gs->seed = J.seed[inst];
    P.gs->pid = J.pid0[inst] + (uint64_t)local;
    P.gs->gamma = J.gamma[inst];
    const cx<double> *f = J.fpm + (int64_t)inst * 30;
    start_values(role, J, P, path);
    if (G::kW == 8) {
        const int j = lane_pose<G>(role);
        const int pidx[6] = {0, j, 8, 8 + j, 15 + j, 22 + j};
#pragma unroll
        for (int p = 0; p < 6; ++p) P.par[6 + p] = f[pidx[p]] - P.par[p];
    } else {
        lds_sync();
        P.par[2 * role + 1] = f[role] - P.par[2 * role];
        if (role + 16 < 30) P.par[2 * (role + 16) + 1] = f[role + 16] - P.par[2 * (role + 16)];
        lds_sync();
    }
    // TrackPath reads the same stream from draw 0 (frg.cu:54,57)
    draw_vec(role, P, 0, o.patch_lo, o.patch_hi, P.u);
    switch_patch<G>(role, P.u, P.Yc);
    {
        cx<double> v[G::kS];
        draw_vec(role, P, 46, o.patch_lo, o.patch_hi, v);
#pragma unroll
        for (int k = 0; k < G::kS; ++k) P.vc[k] = v[k];
    }
    P.gs->draw = 92;
    P.phase = PH_INIT;
    P.stage = 0;
    P.gs->code = 0;
    P.done = false;
    P.s = o.sinit;
    P.tstep = false;
    P.gs->streak = o.caution_steps - 1;
    P.gs->step = 0;
    P.gs->ministep = 0;
    P.gs->err = 0.0;
    P.gs->scale = 1.0;
    P.residual = 100.0;
    P.gs->attempt = 0;
    P.its = 0;
    P.tgoal = 0.0;
    P.gs->st_rk = P.gs->st_newton_its = P.gs->st_redraws = P.gs->st_tan = 0;
    P.gs->st_fp32 = 0;
    P.gs->rec = -1;
    P.pending = false;
    P.gs->rej = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int k = 0; k < G::kS; ++k) P.K[m][k] = mk(0.0, 0.0);
}

// ------------------------------------------------------------------------------------------
// State-machine transitions (htpytrack.cu:83-216).  Each returns with the next solve set up or
// with P.done set.
// ------------------------------------------------------------------------------------------
// A stage point is Yc + s * sum_m kA[stage][m] k_m over all seven slots, with zero coefficients
// for the stages not computed yet in this attempt (DormandPrince65 sums only the earlier ones,
// rungekutta.cu:43-104).  Clearing those slots when an attempt starts keeps a stale non-finite
// value (the YN of a failed corrector, a stage of a rejected attempt) out of the sum: 0 * NaN is
// NaN.
template <int kFirst, class G>
__device__ __forceinline__ void clear_stages(PathState<G> &P) {
#pragma unroll
    for (int m = kFirst; m < 8; ++m)
#pragma unroll
        for (int k = 0; k < G::kS; ++k) P.K[m][k] = mk(0.0, 0.0);
}
template <class G>
__device__ __forceinline__ void start_rk(PathState<G> &P) {
    P.phase = PH_RK;
    P.stage = 0;
    clear_stages<0>(P);
}
template <class G>
__device__ __forceinline__ void main_loop_top(const TrackOpts &o, PathState<G> &P) {
    if (P.gs->step == o.max_steps) {  // :85
        P.gs->code = 1;
        P.done = true;
        return;
    }
    start_rk(P);
}
template <class G>
__device__ __forceinline__ void start_newton(const TrackOpts &o, PathState<G> &P, double tgoal, bool sharp) {
    // YN = Yn (K[1] -> K[0])
#pragma unroll
    for (int k = 0; k < G::kS; ++k) P.K[0][k] = P.K[1][k];
    P.its = 0;
    P.residual = 100.0;
    P.tgoal = tgoal;
    P.phase = sharp ? PH_SHARP : PH_NEWTON;
}
template <class G>
__device__ __forceinline__ void redraw_patch(int role, const TrackOpts &o, PathState<G> &P) {
    draw_vec(role, P, P.gs->draw, o.patch_lo, o.patch_hi, P.u);
    P.gs->draw += 46;
    P.gs->st_redraws++;
    switch_patch<G>(role, P.u, P.K[1]);  // only Yn is moved to the new patch (:151, :210)
}
template <class G>
__device__ __forceinline__ void start_sharpen(const TrackOpts &o, PathState<G> &P) {
    P.gs->attempt = 0;
    start_newton(o, P, 1.0, true);
}
// corrector entry (:133-135)
template <class G>
__device__ __forceinline__ void correct(int role, const TrackOpts &o, PathState<G> &P, double tn) {
    if (fabs(1. - tn) < o.sharp_trigger) {
        start_sharpen(o, P);
        return;
    }
    P.gs->attempt = 0;
    start_newton(o, P, tn, false);
}
// Step-size factor of an RK attempt with error estimate err (rungekutta.cu:107-112):
// 0.8 pow(tol/err, 1./6.), at most 5, and 5 for err <= 1e-16 or NaN.
__device__ __forceinline__ double step_factor(double err, double tol) {
    double scale = err > 10e-17 ? 0.8 * pow_sixth(tol / err) : 5.0;
    if (scale > 5) scale = 5;
    return scale;
}
// end of an RK attempt (:94-123)
template <class G>
__device__ __forceinline__ void finish_rk(int role, const TrackOpts &o, PathState<G> &P) {
    P.gs->st_rk++;
    const double s = P.s;
    cx<double> lte[G::kS], yn[G::kS];
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        cx<double> e = mk(0.0, 0.0), b = mk(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (m == 1) continue;
            e = e + (kE[m] * s) * P.K[m][k];
            b = b + (kB[m] * s) * P.K[m][k];
        }
        lte[k] = e;
        yn[k] = P.Yc[k] + b;
    }
    const double err = grp_norm<G>(lte);
    const double scale = step_factor(err, o.rk_tol);
    P.gs->err = err;
    P.gs->scale = scale;
#pragma unroll
    for (int k = 0; k < G::kS; ++k) P.K[1][k] = yn[k];  // Yn (k2 is not used after the attempt)
    if (P.gs->ministep++ == o.max_ministeps) {  // :95
        P.gs->code = 2;
        P.done = true;
        return;
    }
    if (!(err <= o.rk_tol)) {  // rejected (:97-105)
        if (P.s <= o.smin) {
            P.gs->code = 3;
            P.done = true;
            return;
        }
        P.s *= scale;
        if (P.s < o.smin) P.s = o.smin * 0.9;
        P.gs->streak = 0;
        P.gs->rej++;
        // the retry starts from the same point, velocity and patch, so its first stage tangent is
        // the rejected attempt's k1, still in K[0]: go straight to the second stage
        P.phase = PH_RK;
        P.stage = 1;
        clear_stages<1>(P);
#pragma unroll
        for (int k = 0; k < G::kS; ++k) P.K[7][k] = P.K[0][k];  // the second stage's w row is k1
        return;
    }
    const double tn = grp_t<G>(role, P.K[1]);
    if (tn <= 1. - o.nudge_trigger || fabs(1. - tn) < o.sharp_trigger) {  // :115-118
        correct(role, o, P, tn);
        return;
    }
    P.tstep = true;  // land on t = 1 (:121-123)
    P.s = 1 - grp_t<G>(role, P.Yc);
    start_rk(P);
}
// Features of the step proposal made at an accepted step (see frg_mlp.h), identical in every
// lane of the group.
template <class G>
__device__ __forceinline__ void step_features(int role, const PathState<G> &P, float f[kMlpIn]) {
    const double t = grp_t<G>(role, P.Yc);
    const double ynorm = grp_norm<G>(P.Yc);
    cx<double> v[G::kS];
#pragma unroll
    for (int k = 0; k < G::kS; ++k) v[k] = P.vc[k];
    const double vt = fabs(grp_bcast<G::kW>(v[G::kTSlot].re, role, G::kTLane));
    auto lg = [](double x) { return __logf(fminf(fmaxf((float)x, 1e-30f), 1e30f)); };
    f[0] = (float)t;
    f[1] = lg(P.s);
    f[2] = lg(P.gs->err);
    f[3] = (float)P.its;
    f[4] = lg(P.residual);
    f[5] = lg(ynorm);
    f[6] = -lg(vt);
    f[7] = (float)P.gs->rej;
}

// Step size bounds and the next RK attempt (htpytrack.cu:178-183, then the loop top).
template <class G>
__device__ __forceinline__ void set_step(const TrackOpts &o, PathState<G> &P, double s_next) {
    P.s = s_next;
    if (P.s > o.smax) P.s = o.smax;
    else if (P.s < o.smin) P.s = o.smin * 0.9;
    main_loop_top(o, P);
}

// Accepted step (htpytrack.cu:176-183).  The next step size comes from the classical rule
// s *= scale (scale = 0.8 (tol/err)^(1/6), capped at 5, at most 1 within cautionSteps of a
// rejection), or -- learned control, arclength steps -- from the model, with the same cap and
// caution rule and the classical value as the fallback for a non-finite proposal.  A learned
// proposal is left pending (features in the group's LDS area `xarea`, one per lane 0..7): the
// wave evaluates the model for all its pending groups at once and finish_proposal completes the
// step.
template <class G>
__device__ __forceinline__ void accept_step(int role, const Job &J, PathState<G> &P, float *xarea) {
    const TrackOpts &o = J.o;
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        P.Yc[k] = P.K[0][k];
        P.vc[k] = P.K[7][k];
    }
    const bool use_model = (J.learned || J.steplog) && !P.tstep;
    float f[kMlpIn];
    if (use_model) step_features(role, P, f);
    if (J.steplog && role == 0) {
        // outcome of the previous proposal: the step that would have met the tolerance now
        if (P.gs->rec >= 0 && !P.tstep && P.gs->err > 0.0)
            J.steplog[(int64_t)P.gs->rec * kStepRecord + kMlpIn] = (float)(log(P.s) + log(o.rk_tol / P.gs->err) / 6.0);
        P.gs->rec = -1;
        if (use_model) {
            const unsigned long long i = atomicAdd(J.steplog_count, 1ull);
            if ((int64_t)i < J.steplog_cap) {
                float *rec = J.steplog + (int64_t)i * kStepRecord;
#pragma unroll
                for (int q = 0; q < kMlpIn; ++q) rec[q] = f[q];
                rec[kMlpIn] = __int_as_float(0x7fc00000);  // no outcome yet
                P.gs->rec = (int)i;
            }
        }
    }
    P.gs->streak++;
    if (P.gs->streak < o.caution_steps && P.gs->scale > 1.) P.gs->scale = 1.;
    const double s_classic = P.s * P.gs->scale;
    P.gs->rej = 0;
    P.gs->step++;
    if (J.learned && use_model) {
        float fr = f[0], mu = J.model.xmu[0], sd = J.model.xsd[0];
#pragma unroll
        for (int q = 1; q < kMlpIn; ++q) {
            fr = role == q ? f[q] : fr;
            mu = role == q ? J.model.xmu[q] : mu;
            sd = role == q ? J.model.xsd[q] : sd;
        }
        lds_sync();
        if (role < kMlpIn) xarea[role] = (fr - mu) / sd;
        P.s_classic = s_classic;
        P.pending = true;
        return;
    }
    set_step(o, P, s_classic);
}

// Completes a pending learned proposal with the model output y (log step size).
template <class G>
__device__ __forceinline__ void finish_proposal(const TrackOpts &o, PathState<G> &P, float y) {
    double s_next = P.s_classic;
    // safety factor on the model's step (default 0.6: the minimum of RK attempts per path in a
    // sweep over 0.35..1.15 on 2^18 paths of a seeded task; fewer rejections outweigh the shorter
    // steps) and cap on its growth (default 5, the classical rule's)
    double s_ml = (double)__expf(y) * o.ml_safety;
    if (s_ml == s_ml && s_ml < 1e300) {
        s_ml = fmin(s_ml, o.ml_growth * P.s);
        if (P.gs->streak < o.caution_steps) s_ml = fmin(s_ml, P.s);
        s_next = s_ml;
    }
    P.pending = false;
    set_step(o, P, s_next);
}

// end of a corrector Newton call (:140-164)
template <class G>
__device__ __forceinline__ void finish_newton(int role, const Job &J, PathState<G> &P, float *xarea) {
    const TrackOpts &o = J.o;
    if (P.gs->ministep++ == o.max_ministeps) {
        P.gs->code = 2;
        P.done = true;
        return;
    }
    if (P.residual < o.newt_tol) {
        accept_step(role, J, P, xarea);
        return;
    }
    P.gs->rej++;
    redraw_patch(role, o, P);
    P.gs->attempt++;
    if (P.gs->attempt < o.newt_attempts) {
        start_newton(o, P, grp_t<G>(role, P.K[1]), false);
        return;
    }
    if (P.s <= o.smin) {  // :159
        P.gs->code = 4;
        P.done = true;
        return;
    }
    P.s *= 0.5;
    if (P.s < o.smin) P.s = o.smin * 0.9;
    P.gs->streak = 0;
    start_rk(P);
}

// The homogeneous end point (before dehomogenisation) for callers that asked for it.
template <class G>
__device__ __forceinline__ void store_yfinal(int role, const Job &J, int64_t path, const cx<double> Y[G::kS]) {
    if (!J.yfinal) return;
    cx<double> *yo = J.yfinal + path * 23;
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        const int yi = G::slot_index(role, k);
        if (yi >= 0) yo[yi] = Y[k];
    }
}

// end of a sharpening Newton call (:198-216)
template <class G>
__device__ __forceinline__ void finish_sharp(int role, const Job &J, PathState<G> &P) {
    const TrackOpts &o = J.o;
    if (P.gs->ministep++ == o.max_ministeps) {
        P.gs->code = 2;
        P.done = true;
        return;
    }
    if (P.residual < o.sharp_tol) {
        P.gs->code = 0;
        P.done = true;
        return;
    }
    redraw_patch(role, o, P);
    P.gs->attempt++;
    if (P.gs->attempt < o.sharp_attempts) {
        start_newton(o, P, 1.0, true);
        return;
    }
    P.gs->code = 5;
    P.done = true;
}

// ------------------------------------------------------------------------------------------
// Output of the tracker: dehomogenise (frg.cu:60) and store the 21 values with the code and the
// counters.  The refinement kernel below then polishes code-0 roots and applies CognateSolns.
// ------------------------------------------------------------------------------------------
template <class G>
__device__ __forceinline__ void finish_path(int role, const Job &J, PathState<G> &P) {
    // output point: Yc for codes 1-4 and code 2 before sharpening, the sharpened point otherwise
    const bool from_sharp = (P.phase == PH_SHARP);
    cx<double> Y[G::kS];
#pragma unroll
    for (int k = 0; k < G::kS; ++k) Y[k] = sel(from_sharp, P.K[0][k], P.Yc[k]);
    const int64_t path = P.gs->path;
    store_yfinal<G>(role, J, path, Y);
    // dehomogenise: divide z and Z0 by Z0
    const cx<double> iZ0 = inv(grp_bcast<G::kW>(Y[G::kZSlot], role, G::kZLane));
    cx<double> *out = J.fpt + path * 21;
#pragma unroll
    for (int k = 0; k < G::kS; ++k) {
        const int yi = G::slot_index(role, k);
        if (yi >= 0 && yi < 21) out[yi] = Y[k] * iZ0;
    }
    if (role == 0) {
        J.err[path] = P.gs->code;
        if (J.stats) {
            int *st = J.stats + path * 8;
            st[0] = P.gs->step;
            st[1] = P.gs->st_rk;
            st[2] = P.gs->ministep - P.gs->st_rk;
            st[3] = P.gs->st_newton_its;
            st[4] = P.gs->ministep;
            st[5] = P.gs->st_redraws;
            st[6] = P.gs->st_tan;
            st[7] = P.gs->st_fp32;
        }
    }
}

// ------------------------------------------------------------------------------------------
// One solve for the current phase, then the transition.  8 lanes: FP64 or FP32 (R); 16 lanes:
// FP64.
// ------------------------------------------------------------------------------------------
template <class G, typename R>
__device__ __forceinline__ void micro_step(int role, const Job &J, PathState<G> &P, void *lds) {
    const TrackOpts &o = J.o;
    // evaluation point, w row and right-hand side of the current phase
    const bool newton = (P.phase == PH_NEWTON || P.phase == PH_SHARP);
    const bool rk = (P.phase == PH_RK);
    cx<double> Ye[G::kS], W[G::kS];
    double w_rhs = 1.0;
    {
        // The stage row is read for every phase (stage & 7 is a valid row) and the point chosen
        // afterwards: a read under the phase condition becomes one branch and one wait for the
        // load per coefficient, seven round trips to the cache instead of one.
        double c[7];
        const double s = P.s;
        const double *arow = kA[P.stage & 7];
#pragma unroll
        for (int m = 0; m < 7; ++m) c[m] = arow[m] * s;
        const bool w_last = rk && P.stage > 0;
#pragma unroll
        for (int k = 0; k < G::kS; ++k) {
            cx<double> acc = P.Yc[k];
#pragma unroll
            for (int m = 0; m < 7; ++m) acc = acc + c[m] * P.K[m][k];
            // initial tangent: the current point itself
            Ye[k] = sel(newton, P.K[0][k], sel(rk, acc, P.Yc[k]));
            // w: Vc at init and stage 0, k_{stage-1} (kept in K[7] until stage 8 is stored) for
            // the later stages, e_t for Newton
            W[k] = sel(newton, mk(is_t_slot<G>(role, k) ? 1.0 : 0.0, 0.0), sel(w_last, P.K[7][k], P.vc[k]));
        }
    }
    double tcur = 0.0;
    if (newton) {
        tcur = grp_t<G>(role, Ye);
        w_rhs = P.tgoal - tcur;
    }
    cx<double> x[G::kS];
    if constexpr (G::kW == 16) {
        group_solve16(role, static_cast<GroupLds16 *>(lds), Ye, P.u, W, newton, w_rhs, P.par, P.gs->gamma, x);
    } else {
        GroupLds<R> *L = static_cast<GroupLds<R> *>(lds);
        // FP64 residual of this lane's equations (Newton with FP32 rows only)
        cx<double> fN[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
        if (newton && sizeof(R) != sizeof(double)) {
            // gather the point in FP64 through LDS (two cells per value)
            double *dcell = reinterpret_cast<double *>(&L->phat[0][0]);
            lds_sync();
            dcell[4 * role + 0] = Ye[0].re;
            dcell[4 * role + 1] = Ye[0].im;
            dcell[4 * role + 2] = Ye[1].re;
            dcell[4 * role + 3] = Ye[1].im;
            lds_sync();
            PoseEval<double> e;
#pragma unroll
            for (int i = 0; i < 14; ++i) e.L[i] = mk(dcell[2 * i], dcell[2 * i + 1]);
            e.Z0 = mk(dcell[28], dcell[29]);
            e.t = dcell[30];
            lds_sync();
            e.U = role < 7 ? Ye[2] : mk(1.0, 0.0);
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                e.s[p] = P.par[p];
                e.d[p] = P.par[6 + p];
            }
            e.gamma = P.gs->gamma;
            pose_values(e, fN[0], fN[1], fN[2]);
        }
        group_solve<R>(role, L, Ye, P.u, W, newton, w_rhs, P.par, P.gs->gamma, fN, x);
    }
    if (!newton) {
        P.gs->st_tan++;
        normalize_tangent<G>(role, x, P.tstep && P.phase != PH_INIT);
        if (P.phase == PH_INIT) {
#pragma unroll
            for (int k = 0; k < G::kS; ++k) P.vc[k] = x[k];
            main_loop_top(o, P);
            return;
        }
        // store k_{stage+1}; K[7] holds the latest stage tangent (the next stage's w row) until the
        // eighth stage is stored there
#pragma unroll
        for (int m = 0; m < 7; ++m)
#pragma unroll
            for (int k = 0; k < G::kS; ++k) P.K[m][k] = sel(P.stage == m, x[k], P.K[m][k]);
#pragma unroll
        for (int k = 0; k < G::kS; ++k) P.K[7][k] = x[k];
        P.stage++;
        if (P.stage == 8) finish_rk(role, o, P);
        return;
    }
    // Newton update (htpysupport.cu:475-491)
#pragma unroll
    for (int k = 0; k < G::kS; ++k) P.K[0][k] = P.K[0][k] + x[k];
    P.residual = grp_norm<G>(x);
    P.its++;
    P.gs->st_newton_its++;
    const int max_its = P.phase == PH_SHARP ? o.sharp_its : o.newt_its;
    const double tol = P.phase == PH_SHARP ? o.sharp_tol : o.newt_tol;
    if (P.its < max_its && P.residual > tol) return;  // another iteration
    float *xarea;
    if constexpr (G::kW == 16) xarea = reinterpret_cast<float *>(static_cast<GroupLds16 *>(lds)->a);
    else xarea = reinterpret_cast<float *>(static_cast<GroupLds<double> *>(lds)->a);
    if (P.phase == PH_SHARP) finish_sharp(role, J, P);
    else finish_newton(role, J, P, xarea);
}

// Next path for a persistent group: lane 0 takes a ticket from the device queue.
template <class G>
__device__ __forceinline__ int64_t claim_path(int role, unsigned long long *queue) {
    unsigned long long v = 0;
    if (role == 0) v = atomicAdd(queue, 1ull);
    const int lo = grp_bcast_i<G::kW>((int)(uint32_t)v, role, 0);
    const int hi = grp_bcast_i<G::kW>((int)(uint32_t)(v >> 32), role, 0);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Wave-uniform precision choice for mixed mode: the RK stage tangents are solved in FP32 unless
// some group of the wave needs FP64 (initial tangent, any Newton step -- corrector and
// sharpening -- or a stage within fp64_band of t = 1).
template <class G>
__device__ __forceinline__ bool needs_fp64(int role, const TrackOpts &o, const PathState<G> &P) {
    if (o.precision == 0) return true;
    if (P.phase != PH_RK) return true;
    const double tc = grp_t<G>(role, P.Yc);
    return (1.0 - tc) < o.fp64_band;
}

// ------------------------------------------------------------------------------------------
// Kernels.  Static: group g of the grid tracks path g (the reference's one-thread-per-path
// launch, frg.cu:374).  Persistent: a fixed grid; groups take paths from the queue until it
// runs dry.  With learned step control a group without a path stays in the loop (idle) until
// its whole wave is idle, because the model is evaluated by all 64 lanes of the wave.  The
// 16-lane layout (4 paths per wave) is built for two waves per SIMD: its register budget is 256.
// ------------------------------------------------------------------------------------------
template <bool kPersistent, bool kMixed, class G>
__device__ __forceinline__ void track_body(const Job &J) {
    extern __shared__ unsigned char smem[];
    constexpr int kGroupsPerWave = 64 / G::kW;
    const int role = threadIdx.x & (G::kW - 1);
    const int grp = threadIdx.x / G::kW;
    typename G::Slot *S = reinterpret_cast<typename G::Slot *>(smem) + grp;
    // model scratch: the LDS areas of the wave's groups
    float *xarea0 = reinterpret_cast<float *>((S - (grp % kGroupsPerWave))->lds.a);
    constexpr int kXStride = (int)(sizeof(typename G::Slot) / sizeof(float));
    PathState<G> P;
    if constexpr (G::kW == 8) P.par = S->par[role];
    else P.par = S->par;
    P.vc = S->vc[role];
    P.u = S->u[role];
    P.gs = &S->gs;
    int64_t path = kPersistent ? claim_path<G>(role, J.queue) : (int64_t)blockIdx.x * (blockDim.x / G::kW) + grp;
    bool live = path < J.n_total;
    if (!live && !J.learned) return;
    if (live) init_path(role, J, P, path);
    for (;;) {
        if (live) {
            if (kMixed) {
                const bool need64 = needs_fp64(role, J.o, P);
                if (__any(need64)) {
                    micro_step<G, double>(role, J, P, &S->lds);
                } else {
                    P.gs->st_fp32++;
                    micro_step<G, float>(role, J, P, &S->lds);
                }
            } else {
                micro_step<G, double>(role, J, P, &S->lds);
            }
        }
        if (J.learned) {
            const bool pend = live && P.pending;
            const uint64_t lanes = __ballot(pend && role == 0);
            if (lanes) {
                const float y = mlp_eval_wave<G::kW>(J.model, mlp_lane_weights(J.model), xarea0, kXStride, lanes,
                                                     grp % kGroupsPerWave, role);
                if (pend) finish_proposal(J.o, P, y);
            }
        }
        if (live && P.done) {
            finish_path(role, J, P);
            if (kPersistent) {
                path = claim_path<G>(role, J.queue);
                live = path < J.n_total;
                if (live) init_path(role, J, P, path);
            } else {
                live = false;
            }
        }
        if (J.learned) {
            if (!__any(live)) break;
        } else if (!live) {
            break;
        }
    }
}

template <bool kPersistent, bool kMixed>
__global__ __launch_bounds__(256) void track_kernel(Job J) {
    track_body<kPersistent, kMixed, Lanes8>(J);
}

// kWaves: register budget as resident waves per SIMD (1: up to 512 VGPRs, 2: up to 256).
template <bool kPersistent, int kWaves>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWaves, kWaves))) void track_kernel16(Job J) {
    track_body<kPersistent, false, Lanes16>(J);
}

// ------------------------------------------------------------------------------------------
// Refinement and output, one 8-lane group per path, launched after the tracker:
//  * a code-0 root gets polish_its Newton steps at t = 1 in the dehomogenising patch e_21
//    (frg.cu:44-46) with Z0 = 1: iterative refinement with an extended residual.  The Jacobian
//    rows, the row scaling and the structured solve are the tracker's FP64 Newton step (start
//    values = target, so q = fpm exactly and dF/dt = 0; w = e_t with dt = 0); the residual F is
//    evaluated in double-double at the target and rounded, so the iteration converges to the
//    double nearest the root even when the root is ill-conditioned.  A step is taken only while it
//    is no longer than kPolishCap * max(1, |Y|) (the point is in Newton's quadratic basin);
//  * every path: CognateSolns (frgkinematics.cu:227-285) and the store (frg.cu:63-68).
// ------------------------------------------------------------------------------------------
constexpr double kPolishCap = 1e-6;

// Cognate12 (frgkinematics.cu:185-224): the Stephenson II 1/2 cognate of linkage o[0..6] for
// the coupler point P.
__device__ __forceinline__ void cognate12(const cx<double> o[7], cx<double> P, cx<double> c[7]) {
    const cx<double> A = o[0], B = o[1], C = o[2], D = o[3], F = o[4], G = o[5], H = o[6];
    cx<double> sc = cdiv(B - F, D - F);
    const cx<double> Dp = sc * (D - H) + H, Gp = sc * (G - H) + H;
    sc = cdiv(P - F, H - F);
    c[0] = sc * (A - B) + B;
    c[1] = B;
    c[2] = sc * (C - B) + B;
    c[3] = sc * (Dp - B) + B;
    c[4] = B + (P - F);
    c[5] = sc * (Gp - B) + B;
    c[6] = sc * (H - B) + B;
}

// The 14 linkage values of the group's point, from this lane's slots (lane r < 7 holds z[2r],
// z[2r+1]), gathered through the group's LDS area as doubles.
__device__ __forceinline__ void gather_linkage(int role, GroupLds<double> *L, const cx<double> Y[3],
                                               cx<double> lk[14]) {
    double *dcell = reinterpret_cast<double *>(&L->phat[0][0]);
    lds_sync();
    if (role < 7) {
        dcell[4 * role + 0] = Y[0].re;
        dcell[4 * role + 1] = Y[0].im;
        dcell[4 * role + 2] = Y[1].re;
        dcell[4 * role + 3] = Y[1].im;
    }
    lds_sync();
#pragma unroll
    for (int i = 0; i < 14; ++i) lk[i] = mk(dcell[2 * i], dcell[2 * i + 1]);
    lds_sync();
}

__global__ __launch_bounds__(64) void refine_kernel(Job J) {
    extern __shared__ unsigned char smem[];
    const int role = threadIdx.x & 7;
    const int grp = threadIdx.x >> 3;
    GroupSlot *S = reinterpret_cast<GroupSlot *>(smem) + grp;
    GroupLds<double> *L = &S->lds;
    const int64_t path = (int64_t)blockIdx.x * (blockDim.x >> 3) + grp;
    if (path >= J.n_total) return;  // whole groups
    const int inst = instance_of(J, path);
    const cx<double> *f = J.fpm + (int64_t)inst * 30;
    cx<double> *row = J.fpt + path * 21;
    const int j = role < 7 ? role + 1 : 7;
    // this lane's slots of the dehomogenised point; lane 7 holds Z0 = 1 and t = 1
    cx<double> Y[3];
    if (role < 7) {
        Y[0] = row[2 * role];
        Y[1] = row[2 * role + 1];
        Y[2] = row[14 + role];
    } else {
        Y[0] = mk(1.0, 0.0);
        Y[1] = mk(1.0, 0.0);
        Y[2] = mk(0.0, 0.0);
    }
    if (J.o.polish_its > 0 && J.err[path] == 0) {
        const int pidx[6] = {0, j, 8, 8 + j, 15 + j, 22 + j};  // P0, P_j, Pc0, Pc_j, Q_j, Qc_j
        cx<double> *par = S->par[role];
        cx<dd> qd[6];
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            par[p] = f[pidx[p]];
            par[6 + p] = mk(0.0, 0.0);
            qd[p] = mk(dd(par[p].re), dd(par[p].im));
        }
        cx<double> uu[3], ww[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uu[k] = mk((role == 7 && k == 0) ? 1.0 : 0.0, 0.0);
            ww[k] = mk((role == 7 && k == 1) ? 1.0 : 0.0, 0.0);
        }
        const cx<double> gamma = J.gamma[inst];
        int taken = 0;
        for (int it = 0; it < J.o.polish_its; ++it) {
            cx<double> lk[14];
            gather_linkage(role, L, Y, lk);
            cx<dd> Ld[14];
#pragma unroll
            for (int i = 0; i < 14; ++i) Ld[i] = mk(dd(lk[i].re), dd(lk[i].im));
            const cx<dd> Ud = role < 7 ? mk(dd(Y[2].re), dd(Y[2].im)) : mk(dd(1.0), dd(0.0));
            cx<dd> fI, fII, fIII;
            pose_residuals<dd>(qd, Ld, mk(dd(1.0), dd(0.0)), Ud, fI, fII, fIII);
            // rounded to double: hi is the nearest double to hi + lo
            const cx<double> fN[3] = {mk(fI.re.hi, fI.im.hi), mk(fII.re.hi, fII.im.hi), mk(fIII.re.hi, fIII.im.hi)};
            cx<double> x[3];
            group_solve<double, true>(role, L, Y, uu, ww, true, 0.0, par, gamma, fN, x);
            const double r = grp_norm<Lanes8>(x);
            if (!(r <= kPolishCap * fmax(1.0, grp_norm<Lanes8>(Y)))) break;
#pragma unroll
            for (int k = 0; k < 3; ++k) Y[k] = Y[k] + x[k];
            ++taken;
        }
        if (role == 0 && J.stats && taken) J.stats[path * 8 + 3] += taken;
    }
    // CognateSolns: the root or its 1/2 cognate, whichever has the smaller Re z0
    cx<double> lk[14];
    gather_linkage(role, L, Y, lk);
    const cx<double> P0 = f[0], Pc0 = f[8], Pj = f[j];
    cx<double> c1[7], c2[7];
    cognate12(&lk[0], P0, c1);
    cognate12(&lk[7], Pc0, c2);
    const cx<double> B = lk[1], F = lk[4];
    const cx<double> Udown = cdiv((Pj - B) - Y[2] * (P0 - F), F - B);
    const bool swap = c1[0].re < lk[0].re;
    if (role < 7) {
        cx<double> cog[14];
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            cog[i] = c1[i];
            cog[7 + i] = c2[i];
        }
        cx<double> o0 = Y[0], o1 = Y[1];
#pragma unroll
        for (int i = 0; i < 14; ++i) {
            o0 = sel(swap && i == 2 * role, cog[i], o0);
            o1 = sel(swap && i == 2 * role + 1, cog[i], o1);
        }
        row[2 * role] = o0;
        row[2 * role + 1] = o1;
        row[14 + role] = sel(swap, Udown, Y[2]);
    }
}

hipError_t launch_refine(const Job &job, hipStream_t stream) {
    const int64_t g = (job.n_total + 7) / 8;
    if (g == 0) return hipSuccess;
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)g), dim3(64), 8 * sizeof(GroupSlot), stream, job);
    return hipGetLastError();
}

size_t track_lds_bytes(int block, int lanes) {
    return lanes == 16 ? (size_t)(block / 16) * sizeof(GroupSlot16) : (size_t)(block / 8) * sizeof(GroupSlot);
}

int track_max_groups_per_cu(int block, int kind, int precision, int lanes, int waves) {
    int nb = 0;
    const size_t lds = track_lds_bytes(block, lanes);
    hipError_t e;
    const bool pers = kind == LK_PERSISTENT;
    if (lanes == 16) {
        if (precision) return 0;
        if (waves == 2) {
            e = pers ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel16<true, 2>, block, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel16<false, 2>, block, lds);
        } else {
            e = pers ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel16<true, 1>, block, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel16<false, 1>, block, lds);
        }
    } else if (pers) {
        e = precision ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel<true, true>, block, lds)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel<true, false>, block, lds);
    } else {
        e = precision ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel<false, true>, block, lds)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, track_kernel<false, false>, block, lds);
    }
    if (e != hipSuccess) return 0;
    return nb * (block / lanes);
}

hipError_t launch_track(const Job &job, int kind, int block, int grid, int lanes, int waves, hipStream_t stream) {
    if (block % 64 != 0 || block < 64 || block > 256) return hipErrorInvalidValue;
    if (lanes != 8 && lanes != 16) return hipErrorInvalidValue;
    if (waves != 1 && !(waves == 2 && lanes == 16)) return hipErrorInvalidValue;
    const size_t lds = track_lds_bytes(block, lanes);
    const bool mixed = job.o.precision != 0;
    if (lanes == 16) {
        if (mixed) return hipErrorInvalidValue;  // the 16-lane solve is FP64 only
        const bool pers = kind == LK_PERSISTENT;
        if (waves == 2) {
            if (pers) hipLaunchKernelGGL((track_kernel16<true, 2>), dim3(grid), dim3(block), lds, stream, job);
            else hipLaunchKernelGGL((track_kernel16<false, 2>), dim3(grid), dim3(block), lds, stream, job);
        } else {
            if (pers) hipLaunchKernelGGL((track_kernel16<true, 1>), dim3(grid), dim3(block), lds, stream, job);
            else hipLaunchKernelGGL((track_kernel16<false, 1>), dim3(grid), dim3(block), lds, stream, job);
        }
    } else if (kind == LK_PERSISTENT) {
        if (mixed) hipLaunchKernelGGL((track_kernel<true, true>), dim3(grid), dim3(block), lds, stream, job);
        else hipLaunchKernelGGL((track_kernel<true, false>), dim3(grid), dim3(block), lds, stream, job);
    } else {
        if (mixed) hipLaunchKernelGGL((track_kernel<false, true>), dim3(grid), dim3(block), lds, stream, job);
        else hipLaunchKernelGGL((track_kernel<false, false>), dim3(grid), dim3(block), lds, stream, job);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Stage-level entry points for parity tests: one solve per path on explicit inputs, and the
// start system alone.  They run the same device code as the tracker.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void load_slots(int role, const cx<double> *v23, cx<double> s[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int yi = slot_index(role, k);
        s[k] = yi < 0 ? mk(0.0, 0.0) : v23[yi];
    }
}
__device__ __forceinline__ void store_slots(int role, cx<double> *v23, const cx<double> s[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int yi = slot_index(role, k);
        if (yi >= 0) v23[yi] = s[k];
    }
}

// mode 0: arclength tangent (IVP tstepQ = false), 1: t-step tangent, 2: Newton step to tgoal
// (returns the update), precision 0: FP64, 1: FP32 Jacobian and solve.
template <typename R>
__global__ __launch_bounds__(64) void debug_solve_kernel(DebugSolveArgs A) {
    extern __shared__ unsigned char smem[];
    const int role = threadIdx.x & 7;
    const int grp = threadIdx.x >> 3;
    GroupSlot *S = reinterpret_cast<GroupSlot *>(smem) + grp;
    GroupLds<R> *L = reinterpret_cast<GroupLds<R> *>(&S->lds);
    cx<double> *par = S->par[role];
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 3) + grp;
    if (p >= A.n) return;
    cx<double> Y[3], U[3], W[3], x[3], s6[6], d6[6];
    load_slots(role, A.Y + p * 23, Y);
    load_slots(role, A.u + p * 23, U);
    load_slots(role, A.w + p * 23, W);
    const int j = role < 7 ? role + 1 : 7;
    const cx<double> *sp = A.spm + p * 30;
    const int idx[6] = {0, j, 8, 8 + j, 15 + j, 22 + j};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        s6[q] = sp[idx[q]];
        d6[q] = A.fpm[idx[q]] - s6[q];
        par[q] = s6[q];
        par[6 + q] = d6[q];
    }
    const bool newton = A.mode == 2;
    double w_rhs = 1.0;
    cx<double> fN[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
    if (newton) {
#pragma unroll
        for (int k = 0; k < 3; ++k) W[k] = mk((role == 7 && k == 1) ? 1.0 : 0.0, 0.0);
        w_rhs = A.tgoal[p] - grp_t<Lanes8>(role, Y);
        // FP64 residual values of this lane's equations
        PoseEval<double> e;
        const cx<double> *y = A.Y + p * 23;
#pragma unroll
        for (int i = 0; i < 14; ++i) e.L[i] = y[i];
        e.Z0 = y[21];
        e.t = y[22].re;
        e.U = role < 7 ? Y[2] : mk(1.0, 0.0);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            e.s[q] = s6[q];
            e.d[q] = d6[q];
        }
        e.gamma = A.gamma;
        pose_values(e, fN[0], fN[1], fN[2]);
    }
    group_solve<R>(role, L, Y, U, W, newton, w_rhs, par, A.gamma, fN, x);
    if (!newton) normalize_tangent<Lanes8>(role, x, A.mode == 1);
    store_slots(role, A.x + p * 23, x);
}

// The same with one path per 16-lane group (FP64).
__global__ __launch_bounds__(64) void debug_solve16_kernel(DebugSolveArgs A) {
    extern __shared__ unsigned char smem[];
    const int role = threadIdx.x & 15;
    const int grp = threadIdx.x >> 4;
    GroupSlot16 *S = reinterpret_cast<GroupSlot16 *>(smem) + grp;
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 4) + grp;
    if (p >= A.n) return;
    const cx<double> *sp = A.spm + p * 30;
    for (int i = role; i < 30; i += 16) {
        S->par[2 * i] = sp[i];
        S->par[2 * i + 1] = A.fpm[i] - sp[i];
    }
    cx<double> Y[2], U[2], W[2], x[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int yi = Lanes16::slot_index(role, k);
        Y[k] = yi < 0 ? mk(0.0, 0.0) : A.Y[p * 23 + yi];
        U[k] = yi < 0 ? mk(0.0, 0.0) : A.u[p * 23 + yi];
        W[k] = yi < 0 ? mk(0.0, 0.0) : A.w[p * 23 + yi];
    }
    lds_sync();
    const bool newton = A.mode == 2;
    double w_rhs = 1.0;
    if (newton) {
#pragma unroll
        for (int k = 0; k < 2; ++k) W[k] = mk(is_t_slot<Lanes16>(role, k) ? 1.0 : 0.0, 0.0);
        w_rhs = A.tgoal[p] - grp_t<Lanes16>(role, Y);
    }
    group_solve16(role, &S->lds, Y, U, W, newton, w_rhs, S->par, A.gamma, x);
    if (!newton) normalize_tangent<Lanes16>(role, x, A.mode == 1);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int yi = Lanes16::slot_index(role, k);
        if (yi >= 0) A.x[p * 23 + yi] = x[k];
    }
}

__global__ __launch_bounds__(64) void debug_start_kernel(DebugStartArgs A) {
    const int role = threadIdx.x & 7;
    const int grp = threadIdx.x >> 3;
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 3) + grp;
    if (p >= A.n) return;
    PathState<Lanes8> P;
    GroupScalars gsbuf;
    P.gs = &gsbuf;
    P.gs->seed = A.seed;
    P.gs->pid = A.pid0 + (uint64_t)p;
    const int j = lane_pose<Lanes8>(role);
    StartPose sp;
    start_pose(A.o, P, j, sp);
    cx<double> *z = A.spt + p * 21;
    cx<double> *q = A.spm + p * 30;
    if (role < 7) {
        z[2 * role] = sp.z[2 * role];
        z[2 * role + 1] = sp.z[2 * role + 1];
        z[14 + role] = sp.U;
        q[j] = sp.q[1];
        q[8 + j] = sp.q[3];
        q[15 + j] = sp.q[4];
        q[22 + j] = sp.q[5];
    }
    if (role == 0) {
        q[0] = sp.q[0];
        q[8] = sp.q[2];
    }
}

hipError_t launch_debug_solve(const DebugSolveArgs &a, int precision, int lanes, hipStream_t stream) {
    const int block = 64;
    if (lanes == 16) {
        if (precision) return hipErrorInvalidValue;
        const int grid = (int)((a.n + 3) / 4);
        if (grid == 0) return hipSuccess;
        hipLaunchKernelGGL(debug_solve16_kernel, dim3(grid), dim3(block), 4 * sizeof(GroupSlot16), stream, a);
        return hipGetLastError();
    }
    const int grid = (int)((a.n + 7) / 8);
    if (grid == 0) return hipSuccess;
    if (precision) {
        hipLaunchKernelGGL(debug_solve_kernel<float>, dim3(grid), dim3(block), 8 * sizeof(GroupSlot), stream, a);
    } else {
        hipLaunchKernelGGL(debug_solve_kernel<double>, dim3(grid), dim3(block), 8 * sizeof(GroupSlot), stream, a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Fixed scratch of the allocation-free batch entry point: the work-queue counter and up to
// kNoAllocMaxInst instance descriptions (first path, gamma, seed, first RNG id), one copy per GPU
// for the life of the process.  Host code reaches it only through the address that
// noalloc_scratch_address() returns for the current device.
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(void **ptr) {
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

// Writes the instance descriptions (kernel arguments) into the scratch layout and clears the
// work-queue counter.
__global__ void stage_instances_kernel(NoAllocStage a, unsigned char *base) {
    const int i = threadIdx.x;
    NoAllocLayout L = noalloc_layout(base);
    if (i < a.n_inst) {
        L.gamma[i] = a.gamma[i];
        L.seed[i] = a.seed[i];
        L.pid0[i] = a.pid0[i];
    }
    if (i <= a.n_inst) L.inst_first[i] = a.inst_first[i];
    if (i == 0) *L.queue = 0ull;
}
hipError_t launch_stage_instances(const NoAllocStage &a, void *base, hipStream_t stream) {
    if (a.n_inst < 1 || a.n_inst > kNoAllocMaxInst) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stage_instances_kernel, dim3(1), dim3(kNoAllocMaxInst + 1), 0, stream, a,
                       static_cast<unsigned char *>(base));
    return hipGetLastError();
}

__global__ void reset_counter_kernel(unsigned long long *p) {
    if (threadIdx.x == 0) *p = 0ull;
}
hipError_t launch_reset_counter(unsigned long long *p, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, p);
    return hipGetLastError();
}

__global__ void debug_step_factor_kernel(const double *err, int64_t n, double tol, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = step_factor(err[i], tol);
}
hipError_t launch_debug_step_factor(const double *err, int64_t n, double tol, double *out, hipStream_t stream) {
    const int64_t g = (n + 255) / 256;
    if (g == 0) return hipSuccess;
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(debug_step_factor_kernel, dim3((unsigned)g), dim3(256), 0, stream, err, n, tol, out);
    return hipGetLastError();
}

hipError_t launch_debug_start(const DebugStartArgs &a, hipStream_t stream) {
    const int grid = (int)((a.n + 7) / 8);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_start_kernel, dim3(grid), dim3(64), 0, stream, a);
    return hipGetLastError();
}

}  // namespace frg
