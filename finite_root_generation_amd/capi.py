"""ctypes binding of include/frg.h (libfrg.so, built in-tree by build.py).

This is the host-side mirror of the reference's round interface: ``track`` is one round of
InitRandStates + Homotopy + copies (frg.cu:356-404); ``sort_dedupe_merge`` and ``completeness``
are the host post-processing of frgsupport.cu.  numpy complex128 arrays are passed directly:
their layout is {double re, im}, the same as frg_c64.  There is no fallback: if libfrg.so is
missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as ct
import os
from typing import Optional

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfrg.so")
NV, NP, NY = 21, 30, 23

START_UNIFORM, START_GAUSSIAN = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
SCHED_STATIC, SCHED_PERSISTENT = 0, 1
STEP_CLASSIC, STEP_LEARNED = 0, 1

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "out of memory",
          -4: "unsupported option", -5: "size beyond the fixed limit of the entry point"}

PATH_STATS_FIELDS = ["steps", "rk_attempts", "newton_calls", "newton_its", "ministeps", "redraws",
                     "tangent_solves", "fp32_solves"]
PATH_STATS_DTYPE = np.dtype([(f, np.int32) for f in PATH_STATS_FIELDS])

# every entry point declared in include/frg.h
EXPORTS = ["frg_default_opts", "frg_create", "frg_destroy", "frg_set_opts", "frg_track", "frg_track_device",
           "frg_track_explicit", "frg_track_batch", "frg_generate_starts", "frg_solve_step",
           "frg_sort_dedupe_merge", "frg_lambertw", "frg_completeness", "frg_strerror", "frg_version",
           "frg_last_kernel_ms", "frg_last_refine_ms", "frg_flops", "frg_collect_step_log", "frg_train_step_model",
           "frg_set_step_model", "frg_step_model_eval", "frg_step_model_grad", "frg_get_opts", "frg_autotune", "frg_sort_dedupe_merge_device", "frg_track_batch_noalloc",
           "frg_step_factor"]
STEP_MODEL_SIZE, STEP_MODEL_WEIGHTS, STEP_RECORD = 4819, 4801, 9


class C64(ct.Structure):
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class Opts(ct.Structure):
    """Mirror of frg_opts."""

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
        ("start_mode", ct.c_int32),
        ("precision", ct.c_int32),
        ("fp64_band", ct.c_double),
        ("schedule", ct.c_int32),
        ("block_size", ct.c_int32),
        ("groups_per_cu", ct.c_int32),
        ("step_control", ct.c_int32),
        ("use_graph", ct.c_int32),
        ("lanes_per_path", ct.c_int32),
        ("polish_its", ct.c_int32),
        ("waves_per_simd", ct.c_int32),
        ("pad0", ct.c_int32),
        ("ml_safety", ct.c_double),
        ("ml_growth", ct.c_double),
    ]


class TuneEntry(ct.Structure):
    """Mirror of frg_tune_entry."""

    _fields_ = [("schedule", ct.c_int32), ("block_size", ct.c_int32), ("groups_per_cu", ct.c_int32),
                ("lanes_per_path", ct.c_int32), ("waves_per_simd", ct.c_int32), ("pad0", ct.c_int32),
                ("kernel_ms", ct.c_double), ("paths_per_s", ct.c_double)]


class Stats(ct.Structure):
    """Mirror of frg_stats."""

    _fields_ = [
        ("n_paths", ct.c_int64), ("code_count", ct.c_int64 * 8),
        ("steps", ct.c_int64), ("rk_attempts", ct.c_int64), ("newton_calls", ct.c_int64),
        ("newton_its", ct.c_int64), ("tangent_solves", ct.c_int64),
        ("kernel_ms", ct.c_double), ("total_ms", ct.c_double),
        ("block_size", ct.c_int32), ("grid_size", ct.c_int32), ("schedule", ct.c_int32), ("precision", ct.c_int32),
        ("refine_ms", ct.c_double), ("lanes_per_path", ct.c_int32), ("waves_per_simd", ct.c_int32),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "code_count"}
        d["code_count"] = list(self.code_count)
        return d


class FlopModel(ct.Structure):
    """Mirror of frg_flop_model."""

    _fields_ = [("pose_rows", ct.c_double), ("pose_values", ct.c_double), ("tangent_solve", ct.c_double),
                ("newton_update", ct.c_double), ("rk_combination", ct.c_double)]


class FrgError(RuntimeError):
    pass


_lib = None


def lib() -> ct.CDLL:
    """Loads libfrg.so; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FrgError(f"{LIB_PATH} is not built (run `python finite_root_generation_amd/build.py`)")
        # torch ships its own HIP runtime under the same soname; loading it first makes this library
        # bind to that one, so a process that uses both holds a single runtime
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ct.CDLL(LIB_PATH)
        P = ct.c_void_p
        L.frg_default_opts.argtypes = [ct.POINTER(Opts)]
        L.frg_create.argtypes = [ct.POINTER(P), ct.c_int, ct.POINTER(Opts)]
        L.frg_destroy.argtypes = [P]
        L.frg_destroy.restype = None
        L.frg_set_opts.argtypes = [P, ct.POINTER(Opts)]
        L.frg_track.argtypes = [P, P, C64, ct.c_uint64, ct.c_uint64, ct.c_int64, P, P, P, ct.POINTER(Stats)]
        L.frg_track_device.argtypes = [P, P, C64, ct.c_uint64, ct.c_uint64, ct.c_int64, P, P, P, P]
        L.frg_track_explicit.argtypes = [P, P, C64, P, P, ct.c_uint64, ct.c_uint64, ct.c_int64, P, P, P, P,
                                         ct.POINTER(Stats)]
        L.frg_track_batch.argtypes = [P, ct.c_int32, P, P, P, P, P, P, P, P, ct.POINTER(Stats)]
        L.frg_generate_starts.argtypes = [P, ct.c_uint64, ct.c_uint64, ct.c_int64, P, P]
        L.frg_solve_step.argtypes = [P, ct.c_int32, ct.c_int32, ct.c_int64, P, P, P, P, P, C64, P, P]
        L.frg_step_factor.argtypes = [P, ct.c_int64, P, P]
        L.frg_sort_dedupe_merge.argtypes = [P, ct.c_int64, P, P, ct.c_int64, ct.c_double, P]
        L.frg_sort_dedupe_merge.restype = ct.c_int64
        L.frg_lambertw.argtypes = [ct.c_double]
        L.frg_lambertw.restype = ct.c_double
        L.frg_completeness.argtypes = [ct.c_int64, ct.c_int64]
        L.frg_completeness.restype = ct.c_double
        L.frg_strerror.argtypes = [ct.c_int]
        L.frg_strerror.restype = ct.c_char_p
        L.frg_version.restype = ct.c_int
        L.frg_last_kernel_ms.argtypes = [P, ct.POINTER(ct.c_double)]
        L.frg_last_refine_ms.argtypes = [P, ct.POINTER(ct.c_double)]
        L.frg_track_batch_noalloc.argtypes = [P, ct.c_int32, P, P, P, P, P, P, P, P, P]
        L.frg_get_opts.argtypes = [P, ct.POINTER(Opts)]
        L.frg_sort_dedupe_merge_device.argtypes = [P, P, ct.c_int64, P, P, ct.c_int64, ct.c_double, P,
                                                   ct.POINTER(ct.c_int64), P]
        L.frg_autotune.argtypes = [P, P, C64, ct.c_uint64, ct.c_uint64, ct.c_int64, P, ct.c_int32,
                                   ct.POINTER(ct.c_int32)]
        L.frg_flops.argtypes = [ct.POINTER(FlopModel)]
        L.frg_collect_step_log.argtypes = [P, P, C64, ct.c_uint64, ct.c_uint64, ct.c_int64, P, ct.c_int64,
                                           ct.POINTER(ct.c_int64)]
        L.frg_train_step_model.argtypes = [P, P, ct.c_int64, ct.c_int32, ct.c_int32, ct.c_float, ct.c_uint64, P, P]
        L.frg_set_step_model.argtypes = [P, P]
        L.frg_step_model_grad.argtypes = [P, P, P, ct.c_int64, P, P]
        L.frg_step_model_eval.argtypes = [P, P, ct.c_int64, P]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise FrgError(f"{what} failed: {lib().frg_strerror(rc).decode()} ({rc})")


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def _c64(z) -> C64:
    z = complex(z)
    return C64(z.real, z.imag)


def default_opts(**kw) -> Opts:
    o = Opts()
    _check(lib().frg_default_opts(ct.byref(o)), "frg_default_opts")
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Tracker:
    """A frg_ctx on one HIP device."""

    def __init__(self, device: int = 0, opts: Optional[Opts] = None, **kw):
        self._h = ct.c_void_p()
        self.opts = opts if opts is not None else default_opts(**kw)
        _check(lib().frg_create(ct.byref(self._h), int(device), ct.byref(self.opts)), "frg_create")

    def close(self):
        if self._h:
            lib().frg_destroy(self._h)
            self._h = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_opts(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.opts, k):
                raise KeyError(k)
            setattr(self.opts, k, v)
        _check(lib().frg_set_opts(self._h, ct.byref(self.opts)), "frg_set_opts")

    def track(self, fpm, gamma, seed: int, n: int, first_path_id: int = 0, want_stats: bool = True):
        fpm = _c(fpm, (NP,))
        fpt = np.zeros((n, NV), np.complex128)
        err = np.zeros(n, np.int32)
        st = np.zeros(n, PATH_STATS_DTYPE) if want_stats else None
        summ = Stats()
        _check(lib().frg_track(self._h, _p(fpm), _c64(gamma), seed, first_path_id, n, _p(fpt), _p(err), _p(st),
                               ct.byref(summ)), "frg_track")
        return fpt, err, st, summ

    def track_explicit(self, fpm, gamma, spt, spm, seed: int, first_path_id: int = 0, want_yfinal: bool = False):
        fpm = _c(fpm, (NP,))
        spt = _c(spt)
        spm = _c(spm)
        n = spt.shape[0]
        assert spt.shape == (n, NV) and spm.shape == (n, NP)
        fpt = np.zeros((n, NV), np.complex128)
        err = np.zeros(n, np.int32)
        st = np.zeros(n, PATH_STATS_DTYPE)
        yf = np.zeros((n, NY), np.complex128) if want_yfinal else None
        summ = Stats()
        _check(lib().frg_track_explicit(self._h, _p(fpm), _c64(gamma), _p(spt), _p(spm), seed, first_path_id, n,
                                        _p(fpt), _p(err), _p(st), _p(yf), ct.byref(summ)), "frg_track_explicit")
        return fpt, err, st, yf, summ

    def track_batch(self, fpms, gammas, seeds, first_path_ids, counts):
        fpms = _c(fpms)
        k = fpms.shape[0]
        assert fpms.shape == (k, NP)
        gam = _c(gammas, (k,))
        seeds = np.ascontiguousarray(seeds, np.uint64)
        pids = np.ascontiguousarray(first_path_ids, np.uint64)
        counts = np.ascontiguousarray(counts, np.int64)
        n = int(counts.sum())
        fpt = np.zeros((n, NV), np.complex128)
        err = np.zeros(n, np.int32)
        st = np.zeros(n, PATH_STATS_DTYPE)
        summ = Stats()
        _check(lib().frg_track_batch(self._h, k, _p(fpms), _p(gam), _p(seeds), _p(pids), _p(counts), _p(fpt), _p(err),
                                     _p(st), ct.byref(summ)), "frg_track_batch")
        return fpt, err, st, summ

    def track_device(self, fpm, gamma, seed: int, n: int, d_fpt: int, d_err: int, d_stats: int = 0,
                     first_path_id: int = 0, stream: int = 0):
        """Device-pointer variant: d_fpt, d_err, d_stats are raw device addresses (ints)."""
        fpm = _c(fpm, (NP,))
        _check(lib().frg_track_device(self._h, _p(fpm), _c64(gamma), seed, first_path_id, n, ct.c_void_p(d_fpt),
                                      ct.c_void_p(d_err), ct.c_void_p(d_stats) if d_stats else None,
                                      ct.c_void_p(stream) if stream else None), "frg_track_device")

    def last_kernel_ms(self) -> float:
        v = ct.c_double()
        _check(lib().frg_last_kernel_ms(self._h, ct.byref(v)), "frg_last_kernel_ms")
        return v.value

    def last_refine_ms(self) -> float:
        v = ct.c_double()
        _check(lib().frg_last_refine_ms(self._h, ct.byref(v)), "frg_last_refine_ms")
        return v.value

    def autotune(self, fpm, gamma, seed: int, n_sample: int, first_path_id: int = 0):
        """Measures the candidate launch shapes and keeps the fastest; returns the measurements
        and the chosen (schedule, block_size, groups_per_cu, lanes_per_path, waves_per_simd)."""
        fpm = _c(fpm, (NP,))
        table = (TuneEntry * 128)()
        k = ct.c_int32()
        _check(lib().frg_autotune(self._h, _p(fpm), _c64(gamma), seed, first_path_id, n_sample, table, 128,
                                  ct.byref(k)), "frg_autotune")
        rows = [{f: getattr(table[i], f) for f, _ in TuneEntry._fields_} for i in range(k.value)]
        o = Opts()
        _check(lib().frg_get_opts(self._h, ct.byref(o)), "frg_get_opts")
        self.opts = o
        return rows, (o.schedule, o.block_size, o.groups_per_cu, o.lanes_per_path, o.waves_per_simd)

    def collect_step_log(self, fpm, gamma, seed: int, n: int, cap: int, first_path_id: int = 0):
        fpm = _c(fpm, (NP,))
        rec = np.zeros((cap, STEP_RECORD), np.float32)
        got = ct.c_int64()
        _check(lib().frg_collect_step_log(self._h, _p(fpm), _c64(gamma), seed, first_path_id, n, _p(rec), cap,
                                          ct.byref(got)), "frg_collect_step_log")
        return rec[:got.value]

    def train_step_model(self, records, steps: int = 2000, batch: int = 4096, lr: float = 2e-3, seed: int = 0):
        rec = np.ascontiguousarray(records, np.float32)
        model = np.zeros(STEP_MODEL_SIZE, np.float32)
        loss = np.zeros(max(steps, 1), np.float32)
        _check(lib().frg_train_step_model(self._h, _p(rec), rec.shape[0], steps, batch, lr, seed, _p(model),
                                          _p(loss)), "frg_train_step_model")
        return model, loss[:steps]

    def set_step_model(self, model):
        m = np.ascontiguousarray(model, np.float32)
        assert m.size == STEP_MODEL_SIZE
        _check(lib().frg_set_step_model(self._h, _p(m)), "frg_set_step_model")

    def step_model_eval(self, features):
        """The tracker's inference (loaded model) on n x 8 features: n predicted log step sizes."""
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 8)
        y = np.zeros(x.shape[0], np.float32)
        _check(lib().frg_step_model_eval(self._h, _p(x), x.shape[0], _p(y)), "frg_step_model_eval")
        return y

    def step_model_grad(self, model, records):
        m = np.ascontiguousarray(model, np.float32)
        rec = np.ascontiguousarray(records, np.float32)
        g = np.zeros(STEP_MODEL_WEIGHTS, np.float32)
        loss = ct.c_float()
        _check(lib().frg_step_model_grad(self._h, _p(m), _p(rec), rec.shape[0], _p(g), ct.byref(loss)),
               "frg_step_model_grad")
        return g, loss.value

    def sort_dedupe_merge_device(self, d_coll: int, m: int, d_fpt: int, d_err: int, n: int, d_out: int,
                                 tol: float = 1e-6, stream: int = 0) -> int:
        """Device pointers in, merged length out (see frg_sort_dedupe_merge_device)."""
        k = ct.c_int64()
        _check(lib().frg_sort_dedupe_merge_device(self._h, d_coll, m, d_fpt, d_err, n, float(tol), d_out,
                                                  ct.byref(k), stream or None), "frg_sort_dedupe_merge_device")
        return k.value

    def generate_starts(self, seed: int, n: int, first_path_id: int = 0):
        spt = np.zeros((n, NV), np.complex128)
        spm = np.zeros((n, NP), np.complex128)
        _check(lib().frg_generate_starts(self._h, seed, first_path_id, n, _p(spt), _p(spm)), "frg_generate_starts")
        return spt, spm

    def step_factor(self, err):
        """The tracker's RK step-size factor for the error estimates `err` (frg_step_factor)."""
        e = np.ascontiguousarray(err, np.float64)
        out = np.zeros(e.shape[0], np.float64)
        _check(lib().frg_step_factor(self._h, e.shape[0], _p(e), _p(out)), "frg_step_factor")
        return out

    def solve_step(self, mode: int, Y, u, w, spm, fpm, gamma, tgoal=None, precision: int = PREC_FP64):
        Y = _c(Y)
        n = Y.shape[0]
        u = _c(u, (n, NY))
        w = _c(w, (n, NY)) if w is not None else None
        spm = _c(spm, (n, NP))
        fpm = _c(fpm, (NP,))
        tg = np.ascontiguousarray(tgoal, np.float64) if tgoal is not None else None
        x = np.zeros((n, NY), np.complex128)
        _check(lib().frg_solve_step(self._h, mode, precision, n, _p(Y), _p(u), _p(w), _p(spm), _p(fpm),
                                    _c64(gamma), _p(tg), _p(x)), "frg_solve_step")
        return x


def sort_dedupe_merge(coll, fpt, err, tol: float = 1e-6):
    coll = np.ascontiguousarray(coll, np.complex128).reshape(-1, NV)
    fpt = np.ascontiguousarray(fpt, np.complex128).reshape(-1, NV)
    err = np.ascontiguousarray(err, np.int32).copy()
    out = np.zeros((coll.shape[0] + fpt.shape[0], NV), np.complex128)
    k = lib().frg_sort_dedupe_merge(_p(coll), coll.shape[0], _p(fpt), _p(err), fpt.shape[0], float(tol), _p(out))
    if k < 0:
        _check(int(k), "frg_sort_dedupe_merge")
    return out[:k], err


def flop_model() -> dict:
    m = FlopModel()
    _check(lib().frg_flops(ct.byref(m)), "frg_flops")
    return {f: getattr(m, f) for f, _ in m._fields_}


def path_flops(stats, model=None) -> float:
    """Algorithmic FLOPs of the tracked paths from their per-path counters."""
    m = model or flop_model()
    return float(stats["tangent_solves"].astype(np.float64).sum() * m["tangent_solve"]
                 + stats["newton_its"].astype(np.float64).sum() * m["newton_update"]
                 + stats["rk_attempts"].astype(np.float64).sum() * m["rk_combination"])


def lambertw(z: float) -> float:
    return lib().frg_lambertw(float(z))


def completeness(successes: int, duplicates: int) -> float:
    return lib().frg_completeness(int(successes), int(duplicates))
