// frg_solve16.h -- the tangent / Newton solve with one path per 16-lane group (4 paths per wave).
//
// Same linear system and the same three-stage structured solve as the 8-lane group_solve
// (htpysupport.cu:250-283, :431-477): the 23 x 23 system [JH (row-scaled); w] with U_j eliminated
// inside pose j, the two patch rows reduced, a 16 x 16 Gauss-Jordan and back-substitution.  What
// changes is the distribution over lanes:
//
//   * lane 2p + h (p < 7) evaluates pose j = p+1, half h: h = 0 the unconjugated side (A..H, P, Q),
//     h = 1 the conjugated one (Ac..Hc, Pc, Qc).  Each equation of the pose has the shape
//         F = Z0 K - W0 W1 / U,   W0 = Z0 X - U Y,   W1 = U Xc - Z0 Yc
//     (the reference's expanded forms, htpykinematics.cu:36-122, multiplied out), which is the
//     same in both halves with (alpha, beta) = (Z0, U) for h = 0 and (U, Z0) for h = 1:
//     W_h = alpha X_h - beta Y_h.  The lanes exchange W_h, compute the derivatives with respect to
//     their own seven linkage variables, and form F, dF/dU_j, dF/dZ0 and dF/dt as pair sums of
//     per-half terms, which are bitwise identical in both lanes (so are the row scale and the
//     pivot choice that depend on them);
//   * lanes 14 and 15 own the projective patch row and the velocity / time row;
//   * Gauss-Jordan holds one row per lane: the two reduced rows of pose p in lanes 2p, 2p+1, the
//     reduced patch rows in lanes 14 and 15; the pivot search is four DPP maxima over the group.
//
// Y slots (2 per lane): lane 2p+h slot 0 = z[7h + p] (linkage), lane 2p slot 1 = U_{p+1}
// (z[14+p]), lane 14 slot 0 = Z0, lane 15 slot 0 = t; the other slots are zero.
#pragma once

#include "frg_system.h"

namespace frg {

struct GroupLds16 {
    cx<double> phat[7][17];  // normalised pivot row of each pose (16 shared columns + rhs)
    cx<double> uw[14];       // U-entries of the (scaled) u row [0..6] and of the w row [7..13]
    cx<double> a[32];        // eval point (16) + U (7) / reduced patch rows (32) / pivot row / solution
};

// Column of the 16 x 16 system held in slot 0 of lane l: linkage z[7h + p], then Z0 (14), t (15).
__device__ __forceinline__ int col16(int l) { return l < 14 ? 7 * (l & 1) + (l >> 1) : l; }

// Per-half context of pose j: the half's own linkage o[0..6] (A B C D F G H of that half), the
// other half's p[0..6], Z0, U, 1/U, the gamma-trick factor kappa, (alpha, beta) of W_h, and the
// half's parameters at t (P0, P_j, Q_j), the other half's Q_j and the directions d*.
struct HalfCtx {
    int h;
    cx<double> Z0, U, iU, kap, al, be;
    cx<double> P0, Pj, Q, Qp, dP0, dPj, dQ;
};

// Shared entries of a row from the half's terms: F = Z0 K - W0 W1 / U, dF/dU = Y W1/U -
// W0 Z0 Yc/U^2, dF/dZ0 = K - (dW0/dZ0 W1 + W0 dW1/dZ0)/U, dF/dt = kappa (Z0 Kd - (Wd0 W1 +
// W0 Wd1)/U), each the pair sum of this half's part.  Wh = alpha X - beta Y, wpu = (the other
// half's W) / U, k and kd: this half's part of K and of its directional derivative.
__device__ __forceinline__ void shared_entries(const HalfCtx &c, cx<double> X, cx<double> Y, cx<double> Wh,
                                               cx<double> wpu, cx<double> k, cx<double> kd, cx<double> xz,
                                               cx<double> yz, cx<double> xd, cx<double> yd, cx<double> &gz,
                                               cx<double> &gt, cx<double> &au, cx<double> &f) {
    const cx<double> f_h = c.Z0 * k - mk(0.5, 0.0) * (Wh * wpu);
    const cx<double> au_h = sel(c.h != 0, -((c.Z0 * c.iU) * Y), Y) * wpu;
    const cx<double> dWz = sel(c.h != 0, -Y, X) + c.al * xz - c.be * yz;
    const cx<double> gz_h = k - dWz * wpu;
    const cx<double> gt_h = c.kap * (c.Z0 * kd - (c.al * xd - c.be * yd) * wpu);
    f = pair_sum(f_h);
    au = pair_sum(au_h);
    gz = pair_sum(gz_h);
    gt = pair_sum(gt_h);
}

// Row of equation I_j, this half: own columns g (derivatives by the half's A..H).
__device__ __forceinline__ void half_row_I(const HalfCtx &c, const cx<double> o[7], const cx<double> p[7],
                                           cx<double> g[7], cx<double> &gz, cx<double> &gt, cx<double> &au,
                                           cx<double> &f) {
    const cx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0), half = mk(0.5, 0.0);
    const cx<double> A = o[0], C = o[2], H = o[6], Ap = p[0], Cp = p[2], Hp = p[6];
    const cx<double> CA = C - A, CAp = Cp - Ap, QQ = c.Q * c.Qp;
    const cx<double> X = c.Pj * c.Z0 - A - CA * c.Q, Y = c.P0 * c.Z0 - H;
    const cx<double> Wh = c.al * X - c.be * Y;
    const cx<double> wpu = pair_swap(Wh) * c.iU;
    // K = sum over the halves of  Ap C - Ap A/2 - Cp H + Hp H/2 + CA CAp Q Qp/2
    const cx<double> k = Ap * C - half * (Ap * A) - Cp * H + half * (Hp * H) + half * ((CA * CAp) * QQ);
    g[0] = c.Z0 * (CAp * (one - QQ)) - (c.al * (c.Q - one)) * wpu;  // A
    g[1] = zero;
    g[2] = c.Z0 * (Ap - Hp + CAp * QQ) + (c.al * c.Q) * wpu;  // C
    g[3] = zero;
    g[4] = zero;
    g[5] = zero;
    g[6] = c.Z0 * (Hp - Cp) - c.be * wpu;  // H
    shared_entries(c, X, Y, Wh, wpu, k, (CA * CAp) * (c.dQ * c.Qp), c.Pj, c.P0, c.Z0 * c.dPj - CA * c.dQ,
                   c.Z0 * c.dP0, gz, gt, au, f);
}

// Row of equation II_j, this half.
__device__ __forceinline__ void half_row_II(const HalfCtx &c, const cx<double> o[7], const cx<double> p[7],
                                            cx<double> g[7], cx<double> &gz, cx<double> &gt, cx<double> &au,
                                            cx<double> &f) {
    const cx<double> zero = mk(0.0, 0.0), half = mk(0.5, 0.0);
    const cx<double> B = o[1], F = o[4], BFp = p[1] - p[4];
    const cx<double> X = c.Pj * c.Z0 - B, Y = c.P0 * c.Z0 - F;
    const cx<double> Wh = c.al * X - c.be * Y;
    const cx<double> wpu = pair_swap(Wh) * c.iU;
    g[0] = zero;
    g[1] = c.Z0 * BFp + c.al * wpu;    // B: dX = -1
    g[2] = zero;
    g[3] = zero;
    g[4] = -(c.Z0 * BFp) - c.be * wpu;  // F: dY = -1
    g[5] = zero;
    g[6] = zero;
    shared_entries(c, X, Y, Wh, wpu, half * ((B - F) * BFp), zero, c.Pj, c.P0, c.Z0 * c.dPj, c.Z0 * c.dP0, gz, gt,
                   au, f);
}

// Row of equation III_j, this half.
__device__ __forceinline__ void half_row_III(const HalfCtx &c, const cx<double> o[7], const cx<double> p[7],
                                             cx<double> g[7], cx<double> &gz, cx<double> &gt, cx<double> &au,
                                             cx<double> &f) {
    const cx<double> one = mk(1.0, 0.0), zero = mk(0.0, 0.0), half = mk(0.5, 0.0);
    const cx<double> A = o[0], B = o[1], C = o[2], D = o[3], F = o[4], G = o[5], H = o[6];
    const cx<double> iFB = inv(F - B), iHC = inv(H - C);
    const cx<double> r1 = (D - B) * iFB, r2 = (G - C) * iHC, CA = C - A;
    const cx<double> PjZ = c.Pj * c.Z0, P0Z = c.P0 * c.Z0;
    const cx<double> m = PjZ - B, n = PjZ - A - CA * c.Q, y1 = P0Z - F, y2 = P0Z - H;
    const cx<double> X = (A - B) + CA * c.Q - r1 * m + r2 * n;
    const cx<double> Y = r2 * y2 - r1 * y1;
    const cx<double> Wh = c.al * X - c.be * Y;
    const cx<double> wpu = pair_swap(Wh) * c.iU;
    const cx<double> GDp = p[5] - p[3];
    const cx<double> omr2 = one - r2, r1m1 = r1 - one, r2m1 = r2 - one;
    const cx<double> miFB = m * iFB, niHC = n * iHC, y1iFB = y1 * iFB, y2iHC = y2 * iHC;
    const cx<double> al = c.al, be = c.be;
    // g = Z0 dK - (alpha dX - beta dY) wpu; K's part (G - D)(Gp - Dp)/2
    g[0] = -((al * ((one - c.Q) * omr2)) * wpu);                                               // A
    g[1] = -((al * (r1m1 * (one - miFB)) + be * (y1iFB * r1m1)) * wpu);                       // B
    g[2] = -((al * (omr2 * (c.Q - niHC)) - be * (y2iHC * r2m1)) * wpu);                       // C
    g[3] = -(c.Z0 * GDp) - ((be * y1iFB - al * miFB) * wpu);                                  // D
    g[4] = -((al * (miFB * r1) - be * (r1 * (one + y1iFB))) * wpu);                           // F
    g[5] = c.Z0 * GDp - ((al * niHC - be * y2iHC) * wpu);                                      // G
    g[6] = -((be * (r2 * (one + y2iHC)) - al * (niHC * r2)) * wpu);                           // H
    const cx<double> dr = r2 - r1, Z0dr = c.Z0 * dr;
    shared_entries(c, X, Y, Wh, wpu, half * ((G - D) * GDp), zero, c.Pj * dr, c.P0 * dr,
                   Z0dr * c.dPj + (CA * omr2) * c.dQ, Z0dr * c.dP0, gz, gt, au, f);
}

// Scale a row by 1/max(|re|,|im|) over its entries (pair maximum: the two halves of the row) and
// produce the Newton right-hand side -F s (htpysupport.cu:260-277, :446-471).
__device__ __forceinline__ void scale_half_row(bool newton, cx<double> g[7], cx<double> &gz, cx<double> &gt,
                                               cx<double> &au, cx<double> f, cx<double> &rhs) {
    double m = fmax(fabs(au.re), fabs(au.im));
    m = fmax(m, fmax(fabs(gz.re), fabs(gz.im)));
    m = fmax(m, fmax(fabs(gt.re), fabs(gt.im)));
#pragma unroll
    for (int k = 0; k < 7; ++k) m = fmax(m, fmax(fabs(g[k].re), fabs(g[k].im)));
    m = pair_max(m);
    const double s = recip(m);
#pragma unroll
    for (int k = 0; k < 7; ++k) g[k] = s * g[k];
    gz = s * gz;
    gt = s * gt;
    au = s * au;
    rhs = sel(newton, mk(-f.re * s, -f.im * s), mk(0.0, 0.0));
}

// A half row: own columns (7), then Z0, t and the right-hand side.
struct HalfRow {
    cx<double> c[10];
};

// Pivoting of the three rows of pose j on |dF/dU_j| (first maximum wins, as LinearSolve): row r
// enters; (p, pau) is the pivot candidate so far; the reduced row is written to red.
__device__ __forceinline__ void pivot_in(const HalfRow &r, cx<double> rau, HalfRow &p, cx<double> &pau, HalfRow &red) {
    const bool swap = abs2(rau) > abs2(pau);
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const cx<double> pk = sel(swap, r.c[k], p.c[k]);
        red.c[k] = sel(swap, p.c[k], r.c[k]);
        p.c[k] = pk;
    }
    const cx<double> rU = sel(swap, pau, rau);
    pau = sel(swap, rau, pau);
    const cx<double> m = rU * inv(pau);
#pragma unroll
    for (int k = 0; k < 10; ++k) red.c[k] = fms(red.c[k], m, p.c[k]);
}

// This half's linkage o[k] = a[2k + h] and the other half's p[k] = a[2k + 1 - h] from the
// broadcast evaluation point; the fence keeps the reads at this place (each equation reads them
// again instead of holding 14 values across the three rows).
__device__ __forceinline__ void load_halves(const GroupLds16 *L, int h, cx<double> o[7], cx<double> p[7]) {
    lds_order();
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        o[k] = L->a[2 * k + h];
        p[k] = L->a[2 * k + 1 - h];
    }
}

// The 16-lane group solve.  Ye, uu, ww: this lane's slots of the evaluation point, the patch u and
// the last row w; newton selects the Newton right-hand side (-H scaled; w rhs = w_rhs) over the
// tangent one (zeros; w rhs = 1).  par: the group's 30 parameters as (start, target - start)
// pairs.  Output: this lane's slots of x.
__device__ __forceinline__ void group_solve16(int l, GroupLds16 *L, const cx<double> Ye[2], const cx<double> uu[2],
                                              const cx<double> ww[2], bool newton, double w_rhs, const cx<double> *par,
                                              cx<double> gamma, cx<double> x[2]) {
    const int h = l & 1, pp = l < 14 ? (l >> 1) : 6, j = pp + 1;
    const cx<double> zero = mk(0.0, 0.0);

    // -------- patch-row scale and residual (group reductions) --------
    double um = fmax(fmax(fabs(uu[0].re), fabs(uu[0].im)), fmax(fabs(uu[1].re), fabs(uu[1].im)));
    cx<double> uy = fma(mk(0.0, 0.0), uu[0], Ye[0]);
    uy = fma(uy, uu[1], Ye[1]);
    um = grp_max<16>(um);
    uy = grp_sum<16>(uy);
    const double usc = 1.0 / um;
    const cx<double> bu = sel(newton, mk(-(uy.re - 1.0) * usc, -uy.im * usc), mk(0.0, 0.0));

    // -------- 1. broadcast the evaluation point: a[0..15] slot 0 of every lane, a[16+p] = U --------
    L->a[l] = Ye[0];
    if (l < 14 && h == 0) L->a[16 + pp] = Ye[1];
    lds_sync();
    const cx<double> Z0 = L->a[14];
    const double t = L->a[15].re;
    const cx<double> U = l < 14 ? L->a[16 + pp] : mk(1.0, 0.0);

    // -------- parameters at t: this half's P0, P_j, Q_j and the other half's Q_j --------
    HalfCtx hc;
    hc.h = h;
    {
        const cx<double> den = mk(gamma.re + t * (1.0 - gamma.re), gamma.im * (1.0 - t));
        const cx<double> id = inv(den);
        const cx<double> tau = t * id;
        hc.kap = gamma * (id * id);
        const int iP0 = 8 * h, iPj = 8 * h + j, iQ = 15 + 7 * h + j, iQp = 15 + 7 * (1 - h) + j;
        hc.dP0 = par[2 * iP0 + 1];
        hc.dPj = par[2 * iPj + 1];
        hc.dQ = par[2 * iQ + 1];
        hc.P0 = fma(par[2 * iP0], tau, hc.dP0);
        hc.Pj = fma(par[2 * iPj], tau, hc.dPj);
        hc.Q = fma(par[2 * iQ], tau, hc.dQ);
        hc.Qp = fma(par[2 * iQp], tau, par[2 * iQp + 1]);
    }
    hc.Z0 = Z0;
    hc.U = U;
    hc.iU = inv(U);
    hc.al = h ? U : Z0;
    hc.be = h ? Z0 : U;

    // -------- 2. pose rows (halves), scaling, pivoting on |dF/dU_j| --------
    // III first (it has the most intermediate values), then I and II enter the pivot search; the
    // half's linkage is read from the broadcast point for each equation
    HalfRow pg, ra, rb;
    cx<double> pau;
    {
        cx<double> o[7], q[7], g[7], gz, gt, au, f, rhs;
        load_halves(L, h, o, q);
        half_row_III(hc, o, q, g, gz, gt, au, f);
        scale_half_row(newton, g, gz, gt, au, f, rhs);
#pragma unroll
        for (int k = 0; k < 7; ++k) pg.c[k] = g[k];
        pg.c[7] = gz;
        pg.c[8] = gt;
        pg.c[9] = rhs;
        pau = au;
    }
    {
        cx<double> o[7], q[7], g[7], gz, gt, au, f, rhs;
        load_halves(L, h, o, q);
        half_row_I(hc, o, q, g, gz, gt, au, f);
        scale_half_row(newton, g, gz, gt, au, f, rhs);
        HalfRow r;
#pragma unroll
        for (int k = 0; k < 7; ++k) r.c[k] = g[k];
        r.c[7] = gz;
        r.c[8] = gt;
        r.c[9] = rhs;
        pivot_in(r, au, pg, pau, ra);
    }
    {
        cx<double> o[7], q[7], g[7], gz, gt, au, f, rhs;
        load_halves(L, h, o, q);
        half_row_II(hc, o, q, g, gz, gt, au, f);
        scale_half_row(newton, g, gz, gt, au, f, rhs);
        HalfRow r;
#pragma unroll
        for (int k = 0; k < 7; ++k) r.c[k] = g[k];
        r.c[7] = gz;
        r.c[8] = gt;
        r.c[9] = rhs;
        pivot_in(r, au, pg, pau, rb);
    }
    lds_sync();
    const cx<double> ipau = inv(pau);
    // normalised pivot row -> LDS (each half its columns, the even lane Z0, t and rhs); the
    // U-entries of u (scaled) and w -> LDS
    const cx<double> rho = pg.c[9] * ipau;
    if (l < 14) {
#pragma unroll
        for (int k = 0; k < 7; ++k) L->phat[pp][7 * h + k] = pg.c[k] * ipau;
        if (h == 0) {
            L->phat[pp][14] = pg.c[7] * ipau;
            L->phat[pp][15] = pg.c[8] * ipau;
            L->phat[pp][16] = rho;
            L->uw[pp] = usc * uu[1];
            L->uw[7 + pp] = ww[1];
        }
    }
    lds_sync();

    // -------- 3. eliminate U_1..U_7 from the patch rows; lane l builds column col16(l) --------
    const int cl = col16(l);
    {
        cx<double> cu = usc * uu[0], cw = ww[0];
#pragma unroll
        for (int q7 = 0; q7 < 7; ++q7) {
            const cx<double> a = L->phat[q7][cl];
            cu = fms(cu, L->uw[q7], a);
            cw = fms(cw, L->uw[7 + q7], a);
        }
        lds_sync();
        L->a[cl] = cu;
        L->a[16 + cl] = cw;
    }
    // right-hand sides of the reduced patch rows: b - sum_j (U-entry_j * rho_j)
    const bool pose_lead = l < 14 && h == 0;
    const cx<double> su = grp_sum<16>(sel(pose_lead, (usc * uu[1]) * rho, zero));
    const cx<double> sw = grp_sum<16>(sel(pose_lead, ww[1] * rho, zero));
    lds_sync();

    // -------- the row of this lane in the 16 x 16 system (columns 0..15 and rhs) --------
    // pose lanes: the even lane takes the pose's first reduced row, the odd lane the second; each
    // holds its half of both, so the halves are exchanged within the pair
    cx<double> row[17];
    {
        cx<double> send[7], recv[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            send[k] = sel(h != 0, ra.c[k], rb.c[k]);
            recv[k] = pair_swap(send[k]);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            row[k] = sel(h != 0, recv[k], ra.c[k]);
            row[7 + k] = sel(h != 0, rb.c[k], recv[k]);
        }
        row[14] = sel(h != 0, rb.c[7], ra.c[7]);
        row[15] = sel(h != 0, rb.c[8], ra.c[8]);
        row[16] = sel(h != 0, rb.c[9], ra.c[9]);
    }
    if (l >= 14) {
        const int base = l == 14 ? 0 : 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) row[k] = L->a[base + k];
        row[16] = l == 14 ? bu - su : mk(w_rhs, 0.0) - sw;
    }
    lds_sync();

    // -------- 4. Gauss-Jordan, one row per lane --------
    bool used = false;
    int pc = 0;
    cx<double> ipv = zero;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const uint32_t key = pivot_key(abs2(row[c]), used, l);
        const int gid = 15 - (int)(grp_max_u32<16>(key) & 15u);
        const bool isP = gid == l;
        const cx<double> ip = inv(row[c]);
        if (isP) {
            used = true;
            pc = c;
            ipv = ip;
            L->a[c] = ip;
#pragma unroll
            for (int k = 0; k < 17; ++k)
                if (k > c) L->a[k] = row[k];
        }
        lds_order();
        const cx<double> ipc = L->a[c];
        const cx<double> lm = sel(isP, zero, row[c] * ipc);
#pragma unroll
        for (int k = 0; k < 17; ++k)
            if (k > c) row[k] = fms(row[k], lm, L->a[k]);
        // the next column's pivot lane overwrites L->a only after these reads (LDS requests of a
        // wave complete in issue order; the compiler keeps stores after loads of the same address)
        lds_order();
    }
    lds_sync();
    // -------- 5. solution of the shared unknowns --------
    L->a[pc] = row[16] * ipv;
    lds_sync();
    x[0] = L->a[cl];
    // -------- 6. back-substitute U_j = rho_j - phat_j . x (pair sum of the halves) --------
    cx<double> acc = zero;
#pragma unroll
    for (int k = 0; k < 7; ++k) acc = fma(acc, L->phat[pp][7 * h + k], L->a[7 * h + k]);
    if (h == 0) {
        acc = fma(acc, L->phat[pp][14], L->a[14]);
        acc = fma(acc, L->phat[pp][15], L->a[15]);
    }
    acc = pair_sum(acc);
    x[1] = sel(pose_lead, L->phat[pp][16] - acc, zero);
    lds_sync();
}

}  // namespace frg
