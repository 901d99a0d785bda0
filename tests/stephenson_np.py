"""Independent numpy evaluation of the Stephenson II system, for cross-checking the oracle.

Written from the loop-closure geometry the reference's expanded forms encode
(htpykinematics.cu:36-122): for pose j, with rotation u = U_j (conjugate Z0/U_j) and the
coupler point P_j, each family is  F = Z0*K - (Z0*X - U*Y) * (Xc - (Z0/U)*Yc)  where
  I_j:   X = P_j Z0 - A - (C-A)Q_j,   Y = P0 Z0 - H,
         K = Ac C - Ac A - Hc C + A Cc - H Cc + H Hc + (C-A)(Cc-Ac) Q_j Qc_j
  II_j:  X = P_j Z0 - B,              Y = P0 Z0 - F,   K = (B-F)(Bc-Fc)
  III_j: X = (A-B) + (C-A)Q_j - (D-B)/(F-B) (P_j Z0 - B) + (G-C)/(H-C) (P_j Z0 - A - (C-A)Q_j),
         Y = -(D-B)/(F-B) (P0 Z0 - F) + (G-C)/(H-C) (P0 Z0 - H),   K = (G-D)(Gc-Dc)
and the conjugate quantities Xc, Yc built the same way from the conjugate variables.
Vectorised over a leading batch axis.
"""
import numpy as np


def pose_shapes(Z, q, j):
    """(X, Y, Xc, Yc, K) of equations I_j, II_j, III_j (pose j = 1..7), batch-vectorised; each
    equation is F = Z0*K - (Z0*X - U*Y) * (Xc - (Z0/U)*Yc)."""
    A, B, C, D, F, G, H = (Z[:, i] for i in range(7))
    Ac, Bc, Cc, Dc, Fc, Gc, Hc = (Z[:, i] for i in range(7, 14))
    Z0 = Z[:, 21]
    P0, Pc0 = q[:, 0], q[:, 8]
    Pj, Pcj, Qj, Qcj = q[:, j], q[:, 8 + j], q[:, 15 + j], q[:, 22 + j]
    eq1 = (Pj * Z0 - A - (C - A) * Qj, P0 * Z0 - H, Pcj * Z0 - Ac - (Cc - Ac) * Qcj, Pc0 * Z0 - Hc,
           Ac * C - Ac * A - Hc * C + A * Cc - H * Cc + H * Hc + (C - A) * (Cc - Ac) * Qj * Qcj)
    eq2 = (Pj * Z0 - B, P0 * Z0 - F, Pcj * Z0 - Bc, Pc0 * Z0 - Fc, (B - F) * (Bc - Fc))
    r1, r2 = (D - B) / (F - B), (G - C) / (H - C)
    r1c, r2c = (Dc - Bc) / (Fc - Bc), (Gc - Cc) / (Hc - Cc)
    X = (A - B) + (C - A) * Qj - r1 * (Pj * Z0 - B) + r2 * (Pj * Z0 - A - (C - A) * Qj)
    Y = -r1 * (P0 * Z0 - F) + r2 * (P0 * Z0 - H)
    Xc = (Ac - Bc) + (Cc - Ac) * Qcj - r1c * (Pcj * Z0 - Bc) + r2c * (Pcj * Z0 - Ac - (Cc - Ac) * Qcj)
    Yc = -r1c * (Pc0 * Z0 - Fc) + r2c * (Pc0 * Z0 - Hc)
    eq3 = (X, Y, Xc, Yc, (G - D) * (Gc - Dc))
    return eq1, eq2, eq3


def fsys(Z, q):
    """Residuals F (batch x 21).  Object arrays (mpmath mpc entries) are evaluated in their own
    arithmetic, for extended-precision references."""
    dt = object if (np.asarray(Z).dtype == object or np.asarray(q).dtype == object) else complex
    Z = np.atleast_2d(np.asarray(Z, dt))
    q = np.atleast_2d(np.asarray(q, dt))
    Z0 = Z[:, 21]
    out = np.zeros((Z.shape[0], 21), dt)
    for j in range(1, 8):
        U = Z[:, 13 + j]
        for e, (X, Y, Xc, Yc, K) in enumerate(pose_shapes(Z, q, j)):
            out[:, 7 * e + j - 1] = Z0 * K - (Z0 * X - U * Y) * (Xc - (Z0 / U) * Yc)
    return out


def random_point(rng, n=1, scale=1.0):
    Z = scale * (rng.normal(size=(n, 22)) + 1j * rng.normal(size=(n, 22)))
    q = rng.normal(size=(n, 30)) + 1j * rng.normal(size=(n, 30))
    return Z, q
