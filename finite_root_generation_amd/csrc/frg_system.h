// frg_system.h -- the Stephenson II system for one pose j, evaluated by one lane.
//
// The reference (htpykinematics.cu:8-1384) evaluates FSYS, JVAR (21x22) and JPAR (21x30) as
// expanded, machine-generated expression trees and forms dF/dt = JPAR . dq/dt by a dense 21x30
// product (htpysupport.cu:115-121).  Here the three equations of pose j are evaluated from one
// shared shape (derived from the expanded forms; equal up to rounding):
//
//     F = Z0*K - W*Wc,   W = Z0*X - U*Y,   Wc = Xc - (Z0/U)*Yc
//
//   I_j   X = P_j Z0 - A - (C-A) Q_j,  Y = P0 Z0 - H,   K = Ac(C-A) + Cc(A-H) + Hc(H-C)
//                                                           + (C-A)(Cc-Ac) Q_j Qc_j
//   II_j  X = P_j Z0 - B,              Y = P0 Z0 - F,   K = (B-F)(Bc-Fc)
//   III_j X = (A-B) + (C-A)Q_j - r1 (P_j Z0 - B) + r2 (P_j Z0 - A - (C-A)Q_j),
//         Y = r2 (P0 Z0 - H) - r1 (P0 Z0 - F),  r1 = (D-B)/(F-B), r2 = (G-C)/(H-C),
//         K = (G-D)(Gc-Dc)
//
// (Xc, Yc: the same with conjugate variables and Pc, Qc.)  Derivatives follow by the product
// rule; dF/dt is the directional derivative along dq/dt = gamma/(gamma(1-t)+t)^2 * (f - s), so
// the 21x30 JPAR is never formed.  Only P0, P_j, Pc0, Pc_j, Q_j, Qc_j enter pose j, so a lane
// holds those six start values and six target-minus-start differences.
//
// Row layout: g[0..13] linkage (A B C D F G H Ac Bc Cc Dc Fc Gc Hc), g[14] Z0, g[15] t,
// au = dF/dU_j, f = F.
#pragma once

#include "frg_device.h"

namespace frg {

template <typename R>
struct Row {
    cx<R> g[16];
    cx<R> au;
    cx<R> f;
};

// Parameter slots of one pose: 0 P0, 1 P_j, 2 Pc0, 3 Pc_j, 4 Q_j, 5 Qc_j.
template <typename R>
struct PoseEval {
    // inputs
    cx<R> L[14];
    cx<R> Z0, U;
    R t;
    cx<R> s[6], d[6];  // start values, target - start
    cx<R> gamma;
};

// q(t) = s + t/(gamma(1-t)+t) (f - s) and dq/dt = gamma/(gamma(1-t)+t)^2 (f - s): the reference's
// q and dqdt (htpysupport.cu:37-62) rearranged.  Returns kappa = gamma/den^2.
template <typename R>
__host__ __device__ __forceinline__ cx<R> params_at(const PoseEval<R> &e, cx<R> q[6]) {
    cx<R> den = mk(e.gamma.re + e.t * (R(1) - e.gamma.re), e.gamma.im * (R(1) - e.t));
    cx<R> id = inv(den);
    cx<R> tau = e.t * id;
#pragma unroll
    for (int p = 0; p < 6; ++p) q[p] = fma(e.s[p], tau, e.d[p]);
    return e.gamma * (id * id);
}

// F values of pose j at given parameters q = (P0, P_j, Pc0, Pc_j, Q_j, Qc_j), linkage L[14],
// homogeniser Z0 and rotation U.  Used over double (the Newton residual in FP64 when the
// Jacobian is FP32) and over double-double (the residual of the final root refinement).
template <typename R>
__host__ __device__ __forceinline__ void pose_residuals(const cx<R> q[6], const cx<R> L[14], cx<R> Z0, cx<R> U,
                                                        cx<R> &fI, cx<R> &fII, cx<R> &fIII) {
    const cx<R> P0 = q[0], Pj = q[1], Pc0 = q[2], Pcj = q[3], Q = q[4], Qc = q[5];
    const cx<R> zu = Z0 * inv(U);
    const cx<R> A = L[0], B = L[1], C = L[2], D = L[3], F = L[4], G = L[5], H = L[6];
    const cx<R> Ac = L[7], Bc = L[8], Cc = L[9], Dc = L[10], Fc = L[11], Gc = L[12], Hc = L[13];
    const cx<R> CA = C - A, CcAc = Cc - Ac;
    {
        cx<R> X = Pj * Z0 - A - CA * Q, Y = P0 * Z0 - H;
        cx<R> Xc = Pcj * Z0 - Ac - CcAc * Qc, Yc = Pc0 * Z0 - Hc;
        cx<R> K = Ac * CA + Cc * (A - H) + Hc * (H - C) + (CA * CcAc) * (Q * Qc);
        cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        fI = Z0 * K - W * Wc;
    }
    {
        cx<R> X = Pj * Z0 - B, Y = P0 * Z0 - F, Xc = Pcj * Z0 - Bc, Yc = Pc0 * Z0 - Fc;
        cx<R> K = (B - F) * (Bc - Fc);
        cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        fII = Z0 * K - W * Wc;
    }
    {
        cx<R> iFB = inv(F - B), iHC = inv(H - C), iFBc = inv(Fc - Bc), iHCc = inv(Hc - Cc);
        cx<R> r1 = (D - B) * iFB, r2 = (G - C) * iHC, r1c = (Dc - Bc) * iFBc, r2c = (Gc - Cc) * iHCc;
        cx<R> m = Pj * Z0 - B, n = Pj * Z0 - A - CA * Q, y1 = P0 * Z0 - F, y2 = P0 * Z0 - H;
        cx<R> mc = Pcj * Z0 - Bc, nc = Pcj * Z0 - Ac - CcAc * Qc, y1c = Pc0 * Z0 - Fc, y2c = Pc0 * Z0 - Hc;
        cx<R> X = (A - B) + CA * Q - r1 * m + r2 * n, Y = r2 * y2 - r1 * y1;
        cx<R> Xc = (Ac - Bc) + CcAc * Qc - r1c * mc + r2c * nc, Yc = r2c * y2c - r1c * y1c;
        cx<R> K = (G - D) * (Gc - Dc);
        cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        fIII = Z0 * K - W * Wc;
    }
}

// F values only, at the evaluation point's t.
template <typename R>
__host__ __device__ __forceinline__ void pose_values(const PoseEval<R> &e, cx<R> &fI, cx<R> &fII, cx<R> &fIII) {
    cx<R> q[6];
    params_at(e, q);
    pose_residuals(q, e.L, e.Z0, e.U, fI, fII, fIII);
}

// Quantities shared by the three equations of one pose.
template <typename R>
struct PoseCommon {
    cx<R> P0, Pj, Pc0, Pcj, Q, Qc, dP0, dPj, dPc0, dPcj, dQ, dQc, Z0, U, iU, zu, one, zero;
    cx<R> A, B, C, D, F, G, H, Ac, Bc, Cc, Dc, Fc, Gc, Hc, CA, CcAc, P0Z, PjZ, Pc0Z, PcjZ, kap;
};

template <typename R>
__host__ __device__ __forceinline__ void pose_common(const PoseEval<R> &e, PoseCommon<R> &c) {
    cx<R> q[6];
    c.kap = params_at(e, q);
    c.P0 = q[0], c.Pj = q[1], c.Pc0 = q[2], c.Pcj = q[3], c.Q = q[4], c.Qc = q[5];
    c.dP0 = e.d[0], c.dPj = e.d[1], c.dPc0 = e.d[2], c.dPcj = e.d[3], c.dQ = e.d[4], c.dQc = e.d[5];
    c.Z0 = e.Z0, c.U = e.U;
    c.iU = inv(e.U);
    c.zu = e.Z0 * c.iU;
    c.one = mk(R(1), R(0)), c.zero = mk(R(0), R(0));
    c.A = e.L[0], c.B = e.L[1], c.C = e.L[2], c.D = e.L[3], c.F = e.L[4], c.G = e.L[5], c.H = e.L[6];
    c.Ac = e.L[7], c.Bc = e.L[8], c.Cc = e.L[9], c.Dc = e.L[10], c.Fc = e.L[11], c.Gc = e.L[12], c.Hc = e.L[13];
    c.CA = c.C - c.A, c.CcAc = c.Cc - c.Ac;
    c.P0Z = c.P0 * c.Z0, c.PjZ = c.Pj * c.Z0, c.Pc0Z = c.Pc0 * c.Z0, c.PcjZ = c.Pcj * c.Z0;
}

#define FRG_POSE_UNPACK(c)                                                                            \
    const cx<R> &P0 = c.P0, &Pj = c.Pj, &Pc0 = c.Pc0, &Pcj = c.Pcj, &Q = c.Q, &Qc = c.Qc;             \
    const cx<R> &dP0 = c.dP0, &dPj = c.dPj, &dPc0 = c.dPc0, &dPcj = c.dPcj, &dQ = c.dQ, &dQc = c.dQc; \
    const cx<R> &Z0 = c.Z0, &U = c.U, &iU = c.iU, &zu = c.zu, &one = c.one, &zero = c.zero;           \
    const cx<R> &A = c.A, &B = c.B, &C = c.C, &D = c.D, &F = c.F, &G = c.G, &H = c.H;                 \
    const cx<R> &Ac = c.Ac, &Bc = c.Bc, &Cc = c.Cc, &Dc = c.Dc, &Fc = c.Fc, &Gc = c.Gc, &Hc = c.Hc;   \
    const cx<R> &CA = c.CA, &CcAc = c.CcAc, &P0Z = c.P0Z, &PjZ = c.PjZ, &Pc0Z = c.Pc0Z;               \
    const cx<R> &PcjZ = c.PcjZ, &kap = c.kap;                                                          \
    (void)P0, (void)Pj, (void)Pc0, (void)Pcj, (void)Q, (void)Qc, (void)dP0, (void)dPj, (void)dPc0;      \
    (void)dPcj, (void)dQ, (void)dQc, (void)iU, (void)one, (void)zero, (void)A, (void)B, (void)C, (void)D; \
    (void)F, (void)G, (void)H, (void)Ac, (void)Bc, (void)Cc, (void)Dc, (void)Fc, (void)Gc, (void)Hc;    \
    (void)CA, (void)CcAc, (void)P0Z, (void)PjZ, (void)Pc0Z, (void)PcjZ, (void)kap

// Row of equation I_j, II_j or III_j (EQ = 0, 1, 2) of one pose.
template <int EQ, typename R>
__host__ __device__ __forceinline__ void pose_row(const PoseEval<R> &e, const PoseCommon<R> &c, Row<R> &row) {
    FRG_POSE_UNPACK(c);
    (void)e;
    Row<R> &rI = row, &rII = row, &rIII = row;
    (void)rI, (void)rII, (void)rIII;
    if (EQ == 0)
    // ---------------- I_j ----------------
    {
        const cx<R> QQ = Q * Qc, omQQ = one - QQ, CACc = CA * CcAc;
        const cx<R> X = PjZ - A - CA * Q, Y = P0Z - H;
        const cx<R> Xc = PcjZ - Ac - CcAc * Qc, Yc = Pc0Z - Hc;
        const cx<R> K = Ac * CA + Cc * (A - H) + Hc * (H - C) + CACc * QQ;
        const cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        const cx<R> Z0Wc = Z0 * Wc, UWc = U * Wc;
#pragma unroll
        for (int k = 0; k < 16; ++k) rI.g[k] = zero;
        rI.f = Z0 * K - W * Wc;
        rI.g[0] = Z0 * (CcAc * omQQ) - Z0Wc * (Q - one);               // A
        rI.g[2] = Z0 * (Ac - Hc + CcAc * QQ) + Z0Wc * Q;               // C
        rI.g[6] = Z0 * (Hc - Cc) - UWc;                                // H
        rI.g[7] = Z0 * (CA * omQQ) - W * (Qc - one);                   // Ac
        rI.g[9] = Z0 * (A - H + CA * QQ) + W * Qc;                     // Cc
        rI.g[13] = Z0 * (H - C) - W * zu;                              // Hc
        rI.au = Y * Wc - (W * zu) * (Yc * iU);                         // U_j
        rI.g[14] = K - (X + PjZ - U * P0) * Wc - W * (Pcj - Yc * iU - zu * Pc0);  // Z0
        // t: directional derivative along d = f - s, times kappa
        const cx<R> Xd = Z0 * dPj - CA * dQ, Yd = Z0 * dP0;
        const cx<R> Xcd = Z0 * dPcj - CcAc * dQc, Ycd = Z0 * dPc0;
        const cx<R> Kd = CACc * (dQ * Qc + Q * dQc);
        const cx<R> Wd = Z0 * Xd - U * Yd, Wcd = Xcd - zu * Ycd;
        rI.g[15] = kap * (Z0 * Kd - Wd * Wc - W * Wcd);
    }
    // ---------------- II_j ----------------
    if (EQ == 1) {
        const cx<R> X = PjZ - B, Y = P0Z - F, Xc = PcjZ - Bc, Yc = Pc0Z - Fc;
        const cx<R> BF = B - F, BFc = Bc - Fc;
        const cx<R> K = BF * BFc;
        const cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        const cx<R> Z0Wc = Z0 * Wc;
#pragma unroll
        for (int k = 0; k < 16; ++k) rII.g[k] = zero;
        rII.f = Z0 * K - W * Wc;
        rII.g[1] = Z0 * BFc + Z0Wc;                                    // B
        rII.g[4] = -(Z0 * BFc) - U * Wc;                               // F
        rII.g[8] = Z0 * BF + W;                                        // Bc
        rII.g[11] = -(Z0 * BF) - W * zu;                               // Fc
        rII.au = Y * Wc - (W * zu) * (Yc * iU);
        rII.g[14] = K - (X + PjZ - U * P0) * Wc - W * (Pcj - Yc * iU - zu * Pc0);
        const cx<R> Wd = Z0 * (Z0 * dPj) - U * (Z0 * dP0), Wcd = Z0 * dPcj - zu * (Z0 * dPc0);
        rII.g[15] = kap * (-(Wd * Wc) - W * Wcd);
    }
    // ---------------- III_j ----------------
    if (EQ == 2) {
        const cx<R> iFB = inv(F - B), iHC = inv(H - C), iFBc = inv(Fc - Bc), iHCc = inv(Hc - Cc);
        const cx<R> r1 = (D - B) * iFB, r2 = (G - C) * iHC;
        const cx<R> r1c = (Dc - Bc) * iFBc, r2c = (Gc - Cc) * iHCc;
        const cx<R> m = PjZ - B, n = PjZ - A - CA * Q, y1 = P0Z - F, y2 = P0Z - H;
        const cx<R> mc = PcjZ - Bc, nc = PcjZ - Ac - CcAc * Qc, y1c = Pc0Z - Fc, y2c = Pc0Z - Hc;
        const cx<R> X = (A - B) + CA * Q - r1 * m + r2 * n, Y = r2 * y2 - r1 * y1;
        const cx<R> Xc = (Ac - Bc) + CcAc * Qc - r1c * mc + r2c * nc, Yc = r2c * y2c - r1c * y1c;
        const cx<R> GD = G - D, GDc = Gc - Dc;
        const cx<R> K = GD * GDc;
        const cx<R> W = Z0 * X - U * Y, Wc = Xc - zu * Yc;
        rIII.f = Z0 * K - W * Wc;
        const cx<R> omr2 = one - r2, r1m1 = r1 - one, r2m1 = r2 - one;
        const cx<R> miFB = m * iFB, niHC = n * iHC, y1iFB = y1 * iFB, y2iHC = y2 * iHC;
        // unconjugated: dF = Z0 dK - (Z0 dX - U dY) Wc
        const cx<R> dXA = (one - Q) * omr2;
        const cx<R> dXB = r1m1 * (one - miFB), dYB = -(y1iFB * r1m1);
        const cx<R> dXC = omr2 * (Q - niHC), dYC = y2iHC * r2m1;
        const cx<R> dXD = -miFB, dYD = -y1iFB;
        const cx<R> dXF = miFB * r1, dYF = r1 * (one + y1iFB);
        const cx<R> dXG = niHC, dYG = y2iHC;
        const cx<R> dXH = -(niHC * r2), dYH = -(r2 * (one + y2iHC));
        rIII.g[0] = -((Z0 * dXA) * Wc);
        rIII.g[1] = -((Z0 * dXB - U * dYB) * Wc);
        rIII.g[2] = -((Z0 * dXC - U * dYC) * Wc);
        rIII.g[3] = -(Z0 * GDc) - (Z0 * dXD - U * dYD) * Wc;
        rIII.g[4] = -((Z0 * dXF - U * dYF) * Wc);
        rIII.g[5] = Z0 * GDc - (Z0 * dXG - U * dYG) * Wc;
        rIII.g[6] = -((Z0 * dXH - U * dYH) * Wc);
        // conjugated: dF = Z0 dK - W (dXc - zu dYc)
        const cx<R> omr2c = one - r2c, r1m1c = r1c - one, r2m1c = r2c - one;
        const cx<R> miFBc = mc * iFBc, niHCc = nc * iHCc, y1iFBc = y1c * iFBc, y2iHCc = y2c * iHCc;
        const cx<R> dXAc = (one - Qc) * omr2c;
        const cx<R> dXBc = r1m1c * (one - miFBc), dYBc = -(y1iFBc * r1m1c);
        const cx<R> dXCc = omr2c * (Qc - niHCc), dYCc = y2iHCc * r2m1c;
        const cx<R> dXDc = -miFBc, dYDc = -y1iFBc;
        const cx<R> dXFc = miFBc * r1c, dYFc = r1c * (one + y1iFBc);
        const cx<R> dXGc = niHCc, dYGc = y2iHCc;
        const cx<R> dXHc = -(niHCc * r2c), dYHc = -(r2c * (one + y2iHCc));
        rIII.g[7] = -(W * dXAc);
        rIII.g[8] = -(W * (dXBc - zu * dYBc));
        rIII.g[9] = -(W * (dXCc - zu * dYCc));
        rIII.g[10] = -(Z0 * GD) - W * (dXDc - zu * dYDc);
        rIII.g[11] = -(W * (dXFc - zu * dYFc));
        rIII.g[12] = Z0 * GD - W * (dXGc - zu * dYGc);
        rIII.g[13] = -(W * (dXHc - zu * dYHc));
        rIII.au = Y * Wc - (W * zu) * (Yc * iU);
        const cx<R> dr = r2 - r1, drc = r2c - r1c;
        rIII.g[14] = K - (X + Z0 * (Pj * dr) - U * (P0 * dr)) * Wc - W * (Pcj * drc - Yc * iU - zu * (Pc0 * drc));
        const cx<R> Z0dr = Z0 * dr, Z0drc = Z0 * drc;
        const cx<R> Xd = Z0dr * dPj + (CA * omr2) * dQ, Yd = Z0dr * dP0;
        const cx<R> Xcd = Z0drc * dPcj + (CcAc * omr2c) * dQc, Ycd = Z0drc * dPc0;
        const cx<R> Wd = Z0 * Xd - U * Yd, Wcd = Xcd - zu * Ycd;
        rIII.g[15] = kap * (-(Wd * Wc) - W * Wcd);
    }
}

// Full rows of the three equations of one pose.
template <typename R>
__host__ __device__ __forceinline__ void pose_rows(const PoseEval<R> &e, Row<R> &rI, Row<R> &rII, Row<R> &rIII) {
    PoseCommon<R> c;
    pose_common(e, c);
    pose_row<0>(e, c, rI);
    pose_row<1>(e, c, rII);
    pose_row<2>(e, c, rIII);
}

}  // namespace frg
