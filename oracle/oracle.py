"""ctypes front end of the CPU oracle (oracle/frg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the CPU
baseline leg of ``bench.py``.  The product path never imports it.  numpy complex128 arrays are
passed straight through: their layout is {double re, im}, the same as the oracle's complex type.
"""
from __future__ import annotations

import contextlib
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NV, NP, NY = 21, 30, 23

GAMMA_REF = complex(0.39231132891081666, 1.0062265257923928)  # frg.cu:41


class CCplx(ct.Structure):
    """A complex number passed by value ({double re, im})."""

    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class Opts(ct.Structure):
    """Mirror of otrk_opts (frg_oracle.h)."""

    _fields_ = [
        ("sinit", ct.c_double), ("smax", ct.c_double), ("smin", ct.c_double), ("rk_tol", ct.c_double),
        ("caution_steps", ct.c_int32), ("newt_its", ct.c_int32),
        ("newt_tol", ct.c_double),
        ("newt_attempts", ct.c_int32), ("sharp_its", ct.c_int32),
        ("sharp_tol", ct.c_double),
        ("sharp_attempts", ct.c_int32), ("newt_residual", ct.c_int32),
        ("sharp_trigger", ct.c_double), ("nudge_trigger", ct.c_double),
        ("max_steps", ct.c_int32), ("max_ministeps", ct.c_int32),
        ("patch_lo", ct.c_double), ("patch_hi", ct.c_double),
        ("start_lo", ct.c_double), ("start_hi", ct.c_double),
        ("start_mode", ct.c_int32), ("sharp_residual", ct.c_int32),
        ("polish_its", ct.c_int32), ("pad0", ct.c_int32),
    ]


STATS_FIELDS = ["steps", "rk_attempts", "newton_calls", "newton_its", "ministeps", "redraws",
                "ivp_calls", "end_code", "rejections"]
STATS_DTYPE = np.dtype([(f, np.int32) for f in STATS_FIELDS])

_LIBS: dict = {}


def lib(fast: bool = False) -> ct.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        _LIBS[name] = _configure(ct.CDLL(path))
    return _LIBS[name]


def load(path: str) -> ct.CDLL:
    """Another build of frg_oracle.c (a mutant of the mutation tests), configured like lib()."""
    return _configure(ct.CDLL(path))


@contextlib.contextmanager
def using(L: ct.CDLL):
    """Within the block, the functions of this module call L instead of liboracle.so."""
    prev = _LIBS.get("liboracle.so")
    _LIBS["liboracle.so"] = L
    try:
        yield
    finally:
        if prev is None:
            _LIBS.pop("liboracle.so", None)
        else:
            _LIBS["liboracle.so"] = prev


def _configure(L: ct.CDLL) -> ct.CDLL:
    P = ct.c_void_p
    L.oracle_default_opts.argtypes = [ct.POINTER(Opts)]
    L.oracle_philox4x32_10.argtypes = [P, P, P]
    L.oracle_uniform.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_uint64]
    L.oracle_uniform.restype = ct.c_double
    for fn in ("oracle_cdiv",):
        getattr(L, fn).argtypes = [P, P, P]
    L.oracle_csqrt.argtypes = [P, P]
    L.oracle_refine_root.argtypes = [ct.POINTER(Opts), P, CCplx, P]
    L.oracle_homotopy_explicit_ids.argtypes = [ct.POINTER(Opts), P, CCplx, P, P, ct.c_uint64, P, ct.c_int32,
                                               P, P, ct.c_int]
    for fn in ("oracle_fsys", "oracle_fsys_quad", "oracle_jvar", "oracle_jpar"):
        getattr(L, fn).argtypes = [P, P, P]
    L.oracle_switch_patch.argtypes = [P, P, P]
    for fn in ("oracle_q", "oracle_dqdt"):
        getattr(L, fn).argtypes = [ct.c_double, P, P, P, P]
    for fn in ("oracle_H", "oracle_JH"):
        getattr(L, fn).argtypes = [P, P, P, P, P, P]
    L.oracle_linear_solve.argtypes = [P, P, P]
    L.oracle_linear_solve_newton.argtypes = [P, P, ct.c_double, P]
    L.oracle_ivp.argtypes = [ct.c_int, P, P, P, P, P, P, P]
    L.oracle_newton.argtypes = [ct.c_double, P, P, P, P, P, ct.c_int, ct.c_double, ct.c_int, P, P]
    L.oracle_newton.restype = ct.c_int
    L.oracle_rk65.argtypes = [ct.c_double, ct.c_double, ct.c_int, P, P, P, P, P, P, P, P, P, P]
    L.oracle_track_path.argtypes = [ct.POINTER(Opts), P, P, P, P, ct.c_uint64, ct.c_uint64,
                                    ct.c_uint64, P, P]
    L.oracle_track_path.restype = ct.c_int
    L.oracle_generate_start.argtypes = [ct.POINTER(Opts), ct.c_uint64, ct.c_uint64, P, P]
    L.oracle_cognate12.argtypes = [P, CCplx, P]
    L.oracle_cognate_solns.argtypes = [P, P, P]
    L.oracle_homotopy.argtypes = [ct.POINTER(Opts), P, CCplx, ct.c_uint64, ct.c_uint64,
                                  ct.c_int32, P, P, P, P, ct.c_int]
    L.oracle_homotopy_explicit.argtypes = [ct.POINTER(Opts), P, CCplx, P, P, ct.c_uint64,
                                           ct.c_uint64, ct.c_int32, P, P, P, P, ct.c_int]
    L.oracle_lambertw.argtypes = [ct.c_double]
    L.oracle_lambertw.restype = ct.c_double
    L.oracle_sort_dedupe_merge.argtypes = [P, ct.c_int64, P, P, ct.c_int64, ct.c_double, P]
    L.oracle_sort_dedupe_merge.restype = ct.c_int64
    return L


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().oracle_default_opts(ct.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def _c(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None:
        assert a.size == n, (a.shape, n)
    return a


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _cd(z):
    z = complex(z)
    return (CCplx)(z.real, z.imag)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def uniform(seed, pid, draw):
    return lib().oracle_uniform(seed, pid, draw)


def cdiv(a, b):
    a, b, o = _c([a]), _c([b]), np.zeros(1, np.complex128)
    lib().oracle_cdiv(_p(a), _p(b), _p(o))
    return o[0]


def csqrt(x):
    x, o = _c([x]), np.zeros(1, np.complex128)
    lib().oracle_csqrt(_p(x), _p(o))
    return o[0]


def fsys(Z, q):
    Z, q, o = _c(Z, 22), _c(q, NP), np.zeros(NV, np.complex128)
    lib().oracle_fsys(_p(Z), _p(q), _p(o))
    return o


def homotopy_explicit_ids(opts, fpm, gamma, spt, spm, seed, path_ids, nthreads=0):
    """Explicit starts, path i drawing its patches from RNG id path_ids[i]; returns (fpt, err)."""
    spt = np.ascontiguousarray(spt, np.complex128)
    spm = np.ascontiguousarray(spm, np.complex128)
    n = spt.shape[0]
    ids = np.ascontiguousarray(path_ids, np.uint64)
    assert spt.shape == (n, NV) and spm.shape == (n, NP) and ids.shape == (n,)
    fpt = np.zeros((n, NV), np.complex128)
    err = np.zeros(n, np.int32)
    lib().oracle_homotopy_explicit_ids(ct.byref(opts), _p(_c(fpm, NP)), _cd(gamma), _p(spt), _p(spm), seed,
                                       _p(ids), n, _p(fpt), _p(err), int(nthreads))
    return fpt, err


def refine_root(opts, fpm, gamma, Y):
    """The final refinement of a dehomogenised point Y (23); returns (refined Y, steps taken)."""
    Y = _c(Y, 23).copy()
    k = lib().oracle_refine_root(ct.byref(opts), _p(_c(fpm, 30)), _cd(gamma), _p(Y))
    return Y, k


def fsys_quad(Z, q):
    """FSYS evaluated in IEEE binary128 and rounded to double (the refinement's residual)."""
    Z = _c(Z, 22)
    q = _c(q, 30)
    o = np.zeros(21, complex)
    lib().oracle_fsys_quad(_p(Z), _p(q), _p(o))
    return o


def jvar(Z, q):
    Z, q, o = _c(Z, 22), _c(q, NP), np.zeros((NV, NV + 1), np.complex128)
    lib().oracle_jvar(_p(Z), _p(q), _p(o))
    return o


def jpar(Z, q):
    Z, q, o = _c(Z, 22), _c(q, NP), np.zeros((NV, NP), np.complex128)
    lib().oracle_jpar(_p(Z), _p(q), _p(o))
    return o


def switch_patch(u, Y):
    u, Y, o = _c(u, NY), _c(Y, NY), np.zeros(NY, np.complex128)
    lib().oracle_switch_patch(_p(u), _p(Y), _p(o))
    return o


def q(t, spm, fpm, gamma):
    s, f, g, o = _c(spm, NP), _c(fpm, NP), _c([gamma]), np.zeros(NP, np.complex128)
    lib().oracle_q(t, _p(s), _p(f), _p(g), _p(o))
    return o


def dqdt(t, spm, fpm, gamma):
    s, f, g, o = _c(spm, NP), _c(fpm, NP), _c([gamma]), np.zeros(NP, np.complex128)
    lib().oracle_dqdt(t, _p(s), _p(f), _p(g), _p(o))
    return o


def H(Y, u, spm, fpm, gamma):
    Y, u, s, f, g = _c(Y, NY), _c(u, NY), _c(spm, NP), _c(fpm, NP), _c([gamma])
    o = np.zeros(NV + 1, np.complex128)
    lib().oracle_H(_p(Y), _p(u), _p(s), _p(f), _p(g), _p(o))
    return o


def JH(Y, u, spm, fpm, gamma):
    Y, u, s, f, g = _c(Y, NY), _c(u, NY), _c(spm, NP), _c(fpm, NP), _c([gamma])
    o = np.zeros((NV + 1, NY), np.complex128)
    lib().oracle_JH(_p(Y), _p(u), _p(s), _p(f), _p(g), _p(o))
    return o


def linear_solve(M, b):
    M, b, x = _c(M, NY * NY).copy(), _c(b, NY).copy(), np.zeros(NY, np.complex128)
    lib().oracle_linear_solve(_p(M), _p(b), _p(x))
    return x


def linear_solve_newton(M, b, tdiff):
    M, b, x = _c(M, NY * NY).copy(), _c(b, NV + 1).copy(), np.zeros(NY, np.complex128)
    lib().oracle_linear_solve_newton(_p(M), _p(b), float(tdiff), _p(x))
    return x


def ivp(tstep, Y, V, u, spm, fpm, gamma):
    Y, V, u, s, f, g = _c(Y, NY), _c(V, NY), _c(u, NY), _c(spm, NP), _c(fpm, NP), _c([gamma])
    o = np.zeros(NY, np.complex128)
    lib().oracle_ivp(int(tstep), _p(Y), _p(V), _p(u), _p(s), _p(f), _p(g), _p(o))
    return o


def newton(tgoal, Vn, u, spm, fpm, gamma, max_its, tol, residual_type, Y0):
    Vn, u, s, f, g = _c(Vn, NY), _c(u, NY), _c(spm, NP), _c(fpm, NP), _c([gamma])
    Y = _c(Y0, NY).copy()
    res = ct.c_double()
    its = lib().oracle_newton(float(tgoal), _p(Vn), _p(u), _p(s), _p(f), _p(g), int(max_its), float(tol),
                              int(residual_type), _p(Y), ct.byref(res))
    return Y, res.value, its


def rk65(s, tol, tstep, Yc, Vc, u, spm, fpm, gamma):
    Yc, Vc, u, sp, fp, g = _c(Yc, NY), _c(Vc, NY), _c(u, NY), _c(spm, NP), _c(fpm, NP), _c([gamma])
    Yn, Vn = np.zeros(NY, np.complex128), np.zeros(NY, np.complex128)
    err, scale = ct.c_double(), ct.c_double()
    lib().oracle_rk65(float(s), float(tol), int(tstep), _p(Yc), _p(Vc), _p(u), _p(sp), _p(fp), _p(g),
                      _p(Yn), _p(Vn), ct.byref(err), ct.byref(scale))
    return Yn, Vn, err.value, scale.value


def generate_start(opts, seed, pid):
    z, qv = np.zeros(NV, np.complex128), np.zeros(NP, np.complex128)
    lib().oracle_generate_start(ct.byref(opts), seed, pid, _p(z), _p(qv))
    return z, qv


def cognate12(orig, P):
    o, out = _c(orig, 7), np.zeros(7, np.complex128)
    lib().oracle_cognate12(_p(o), _cd(P), _p(out))
    return out


def cognate_solns(z, qv):
    z, qv, o = _c(z, NV), _c(qv, NP), np.zeros((2, NV), np.complex128)
    lib().oracle_cognate_solns(_p(z), _p(qv), _p(o))
    return o


def track_path(opts, spt, spm, fpm, gamma, seed, pid, first_draw=0):
    spt, spm, fpm, g = _c(spt, NV), _c(spm, NP), _c(fpm, NP), _c([gamma])
    Yf = np.zeros(NY, np.complex128)
    st = np.zeros(1, STATS_DTYPE)
    code = lib().oracle_track_path(ct.byref(opts), _p(spt), _p(spm), _p(fpm), _p(g), seed, pid, first_draw,
                                   _p(Yf), _p(st))
    return code, Yf, st[0]


def homotopy(opts, fpm, gamma, seed, first_path_id, n, nthreads=0, fast=False, want_yfinal=False):
    fpm = _c(fpm, NP)
    fpt = np.zeros((n, NV), np.complex128)
    err = np.zeros(n, np.int32)
    stats = np.zeros(n, STATS_DTYPE)
    yf = np.zeros((n, NY), np.complex128) if want_yfinal else None
    lib(fast).oracle_homotopy(ct.byref(opts), _p(fpm), _cd(gamma), seed, first_path_id, n, _p(fpt), _p(err),
                              _p(stats), _p(yf) if yf is not None else None, int(nthreads))
    return fpt, err, stats, yf


def homotopy_explicit(opts, fpm, gamma, spt, spm, seed, first_path_id, nthreads=0, fast=False,
                      want_yfinal=False):
    spt = np.ascontiguousarray(spt, np.complex128)
    spm = np.ascontiguousarray(spm, np.complex128)
    n = spt.shape[0]
    assert spt.shape == (n, NV) and spm.shape == (n, NP)
    fpm = _c(fpm, NP)
    fpt = np.zeros((n, NV), np.complex128)
    err = np.zeros(n, np.int32)
    stats = np.zeros(n, STATS_DTYPE)
    yf = np.zeros((n, NY), np.complex128) if want_yfinal else None
    lib(fast).oracle_homotopy_explicit(ct.byref(opts), _p(fpm), _cd(gamma), _p(spt), _p(spm), seed,
                                       first_path_id, n, _p(fpt), _p(err), _p(stats),
                                       _p(yf) if yf is not None else None, int(nthreads))
    return fpt, err, stats, yf


def lambertw(z):
    return lib().oracle_lambertw(float(z))


def sort_dedupe_merge(coll, fpt, err, tol=1e-6):
    coll = np.ascontiguousarray(coll, np.complex128).reshape(-1, NV)
    fpt = np.ascontiguousarray(fpt, np.complex128).reshape(-1, NV)
    err = np.ascontiguousarray(err, np.int32).copy()
    out = np.zeros((coll.shape[0] + fpt.shape[0], NV), np.complex128)
    k = lib().oracle_sort_dedupe_merge(_p(coll), coll.shape[0], _p(fpt), _p(err), fpt.shape[0], float(tol),
                                       _p(out))
    return out[:k], err
