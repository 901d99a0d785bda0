// frg_flops.cpp -- algorithmic floating-point operation counts of the tracker's formulation.
//
// The per-pose system evaluation is counted by running the kernel's own templates
// (pose_rows / pose_values from frg_system.h) on the host with an operation-counting scalar:
// each real +, -, *, / counts 1 (a fused multiply-add counts 2, as the two operations it
// replaces).  Sign flips and selects count 0.  The rest of a solve is counted from the algorithm
// as executed once per path (not per lane): row scaling, in-lane elimination of U_j, reduction of
// the two patch rows, an LU-equivalent 16 x 16 solve with partial pivoting (pivot keys
// included), back-substitution of U_1..U_7 and the tangent normalisation or Newton update.
// These are the per-unit figures behind the roofline's algorithmic FLOPs.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/frg.h"
#include "frg_system.h"

namespace {

struct Counter {
    int64_t n = 0;
};
Counter g_cnt;

struct Cnt {
    double v;
    Cnt(double x = 0.0) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { g_cnt.n++; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { g_cnt.n++; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { g_cnt.n++; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { g_cnt.n++; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }

frg::PoseEval<Cnt> sample_pose() {
    frg::PoseEval<Cnt> e;
    for (int k = 0; k < 14; ++k) e.L[k] = frg::mk(Cnt(0.3 + k), Cnt(0.1 * k - 0.2));
    e.Z0 = frg::mk(Cnt(0.9), Cnt(0.1));
    e.U = frg::mk(Cnt(1.1), Cnt(-0.3));
    e.t = Cnt(0.4);
    for (int p = 0; p < 6; ++p) {
        e.s[p] = frg::mk(Cnt(0.5 + p), Cnt(0.2));
        e.d[p] = frg::mk(Cnt(-0.1), Cnt(0.3 * p));
    }
    e.gamma = frg::mk(Cnt(0.4), Cnt(0.9));
    return e;
}

int64_t count_pose_rows() {
    frg::PoseEval<Cnt> e = sample_pose();
    frg::Row<Cnt> a, b, c;
    g_cnt.n = 0;
    frg::pose_rows(e, a, b, c);
    return g_cnt.n;
}

int64_t count_pose_values() {
    frg::PoseEval<Cnt> e = sample_pose();
    frg::cx<Cnt> a, b, c;
    g_cnt.n = 0;
    frg::pose_values(e, a, b, c);
    return g_cnt.n;
}

// complex op costs in real flops
constexpr double kCmul = 6, kCfma = 8, kAbs2 = 3, kCinv = 6;  // inv: |a|^2 (3), 1/x (1), 2 mul

double solve_rest(bool newton) {
    double f = 0;
    // per pose: row scaling (3 reciprocals, 3 x 17 x 2 scalings), Newton rhs scaling (3 x 2),
    // pivot keys (3 |.|^2), pivot reciprocal, 2 multipliers, 2 x 17 fused updates, 17 normalised
    // pivot entries
    const double pose = 3 + 3 * 17 * 2 + (newton ? 6 : 0) + 3 * kAbs2 + kCinv + 2 * kCmul + 2 * 17 * kCfma + 17 * kCmul;
    f += 7 * pose;
    // patch rows: u scale (23 x 2 + 1), u.Y (23 fused), U-elimination of u and w over 16
    // columns (2 x 16 x 7 fused), their right-hand sides (2 x 7 products + 2 x 7 adds)
    f += 23 * 2 + 1 + 23 * kCfma + 2 * 16 * 7 * kCfma + 2 * 7 * kCmul + 2 * 7 * 2;
    // 16 x 16 with one right-hand side: at column c, keys of the 16-c candidates, the pivot
    // reciprocal, and for the 15-c rows below a multiplier and 16-c fused updates
    for (int c = 0; c < 16; ++c) f += (16 - c) * kAbs2 + kCinv + (15 - c) * (kCmul + (16 - c) * kCfma);
    // back substitution
    for (int i = 0; i < 16; ++i) f += (15 - i) * kCfma + kCmul;
    // U_j = rho_j - phat_j . x
    f += 7 * 16 * kCfma;
    // tangent: norm (23 x 4 + sqrt), phase factor (~8), scaling (23 cmul); Newton: Y += x, norm
    f += newton ? (23 * 2 + 23 * 4 + 1) : (23 * 4 + 1 + 8 + 23 * kCmul);
    return f;
}

double rk_combination() {
    // stage points 2..8 (stage i uses i-1 earlier stages; stage 8 uses 6): real*complex fused
    // into complex accumulators (4 flops per term and component), 23 components
    double f = 0;
    const int terms[8] = {0, 1, 2, 3, 4, 5, 6, 6};
    for (int i = 0; i < 8; ++i) f += terms[i] * 23 * 4 + terms[i];
    // error estimate (7 terms), its norm, the step factor, the propagated solution (7 terms)
    f += 7 * 23 * 4 + 23 * 4 + 1 + 4 + 7 * 23 * 4 + 23 * 2;
    return f;
}

}  // namespace

extern "C" int frg_flops(frg_flop_model *out) {
    if (!out) return FRG_EINVAL;
    const double rows = (double)count_pose_rows();
    const double vals = (double)count_pose_values();
    out->pose_rows = rows;
    out->pose_values = vals;
    out->tangent_solve = 7 * rows + solve_rest(false);
    out->newton_update = 7 * rows + solve_rest(true);
    out->rk_combination = rk_combination();
    return FRG_OK;
}
