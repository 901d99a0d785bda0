"""Root comparison shared by the parity tests.

Two roots agree when every one of their 21 dehomogenised, cognate-canonicalised components
differs by at most ROOT_TOL = 1e-8 in absolute value (max |dx| <= 1e-8), after the tracker's FP64
refinement.  The relative figure max |dx| / max(1, |x|) is reported beside it, not tested.
"""
import numpy as np

ROOT_TOL = 1e-8


def abs_diff(a, b):
    """Per-row max |a - b| over the components (rows of 21 complex values)."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    return np.max(np.abs(a - b), axis=1)


def rel_diff(a, b):
    """Per-row max |a - b| / max(1, |b|) (reported only)."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    return np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)), axis=1)


def found_anywhere(roots_a, x, tol=ROOT_TOL):
    """True when some row of roots_a (sorted on Re z0) agrees with root x to tol."""
    keys = roots_a[:, 0].real
    lo = np.searchsorted(keys, x[0].real - tol, side="left")
    hi = np.searchsorted(keys, x[0].real + tol, side="right")
    return hi > lo and bool(np.any(abs_diff(roots_a[lo:hi], x[None]) <= tol))
