"""CPU-only checks of libfrg.so: it loads, exports every entry point include/frg.h declares,
its defaults are the reference's constants, and its host post-processing (duplicate marking,
merge, Lambert W, completeness estimate) agrees with the oracle.  No GPU call is made."""
import os
import re

import numpy as np
import pytest

from finite_root_generation_amd import capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "frg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(frg_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_entry_points()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(capi.EXPORTS) == names


def test_defaults_match_reference_constants():
    o = capi.default_opts()
    r = O.default_opts()
    for f, _ in O.Opts._fields_:
        if f == "pad0":
            continue
        if f == "sharp_residual":  # the ABI has one residual type for both Newton uses
            assert r.sharp_residual == o.newt_residual
            continue
        assert getattr(o, f) == getattr(r, f), f
    assert o.precision == capi.PREC_FP64 and o.schedule == capi.SCHED_STATIC
    assert capi.lib().frg_version() >= 10000


def test_strerror():
    assert capi.lib().frg_strerror(0) == b"ok"
    assert capi.lib().frg_strerror(-4) == b"unsupported option"


def test_counts_beyond_the_per_call_limit_are_rejected():
    """A path or root count above FRG_MAX_COUNT (2^36) is refused before anything is allocated;
    a null context fails the same way for every entry point that takes one."""
    L = capi.lib()
    big = (1 << 36) + 1
    one = np.zeros((1, 21), np.complex128)
    err = np.zeros(1, np.int32)
    out = np.zeros((2, 21), np.complex128)
    assert L.frg_sort_dedupe_merge(capi._p(one), big, capi._p(one), capi._p(err), 1, 1e-6, capi._p(out)) == -1
    assert L.frg_sort_dedupe_merge(capi._p(one), 1, capi._p(one), capi._p(err), big, 1e-6, capi._p(out)) == -1
    other = np.ones((1, 21), np.complex128)
    assert L.frg_sort_dedupe_merge(capi._p(one), 1, capi._p(other), capi._p(err), 1, 1e-6, capi._p(out)) == 2
    assert L.frg_step_factor(None, 1, capi._p(np.ones(1)), capi._p(np.ones(1))) == -1


def _roots(rng, n, base):
    picks = rng.integers(0, base.shape[0], n)
    return base[picks] + 1e-9 * (rng.normal(size=(n, 21)) + 1j * rng.normal(size=(n, 21)))


def test_dedupe_merge_matches_oracle_over_rounds():
    rng = np.random.default_rng(0)
    base = rng.normal(size=(300, 21)) + 1j * rng.normal(size=(300, 21))
    base[::7, 0] = base[::7, 0].real.round(7) + 1j * base[::7, 0].imag  # near-ties in Re z0
    coll_c = np.zeros((0, 21), complex)
    coll_o = np.zeros((0, 21), complex)
    total_dup = 0
    for r in range(4):
        fpt = _roots(rng, 200, base)
        err = rng.choice([0] * 9 + [1, 3, 5], 200).astype(np.int32)
        fpt[err == 5] = np.nan
        coll_c, e_c = capi.sort_dedupe_merge(coll_c, fpt, err)
        coll_o, e_o = O.sort_dedupe_merge(coll_o, fpt, err)
        assert np.array_equal(e_c, e_o), r
        assert np.array_equal(coll_c.view(np.float64), coll_o.view(np.float64))
        total_dup += int((e_c == 6).sum())
    # every distinct base root seen is collected exactly once
    assert len(coll_c) <= base.shape[0]
    assert total_dup > 0
    assert np.all(np.diff(coll_c[:, 0].real) >= 0)


def test_dedupe_empty_inputs():
    out, e = capi.sort_dedupe_merge(np.zeros((0, 21)), np.zeros((0, 21)), np.zeros(0, np.int32))
    assert out.shape == (0, 21) and e.shape == (0,)
    x = np.ones((3, 21), complex)
    out, e = capi.sort_dedupe_merge(np.zeros((0, 21)), x, np.zeros(3, np.int32))
    assert list(e) == [0, 6, 6] and len(out) == 1


def test_lambertw_and_completeness():
    from scipy.special import lambertw
    for z in [-0.3678794, -0.2, -0.01, 0.0, 0.3, 2.0, 50.0]:
        assert abs(capi.lambertw(z) - lambertw(z).real) < 1e-13 * (1 + abs(z)), z
        assert abs(capi.lambertw(z) - O.lambertw(z)) < 1e-12 * (1 + abs(z)), z
    # completeness 1 + r W0(-e^{-1/r}/r) (frg.cu:464-465) against the oracle's Lambert W
    for s, d in [(1000, 1), (1000, 100), (500, 500), (10, 1000)]:
        r = s / (s + d)
        ref = r * O.lambertw((-1 / r) * np.exp(-1 / r)) + 1
        assert abs(capi.completeness(s, d) - ref) < 1e-9
    assert capi.completeness(0, 0) == 0.0
    # sampling with replacement: after k draws from N roots, unique ~ N(1 - e^{-k/N})
    N, k = 10000, 15000
    uniq = N * (1 - np.exp(-k / N))
    est = capi.completeness(int(round(uniq)), int(round(k - uniq)))
    assert abs(est - uniq / N) < 1e-3


def torch_has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_create_without_gpu_fails_loudly():
    if torch_has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.FrgError):
        capi.Tracker(0)


def test_round_driver_is_built_and_fails_loudly_without_gpu(tmp_path):
    """frg_rounds (the C++ round driver over the C ABI) rejects bad arguments and, with no GPU,
    stops at frg_create without writing any file."""
    import subprocess
    exe = os.path.join(os.path.dirname(capi.LIB_PATH), "frg_rounds")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "nRound input error" in r.stderr
    r = subprocess.run([exe, "1", "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    r = subprocess.run([exe, "1", "--learned", "--train-paths", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--train-paths" in r.stderr
    r = subprocess.run([exe, "1", "--rounds-per-launch", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--rounds-per-launch" in r.stderr
    if not torch_has_gpu():
        r = subprocess.run([exe, "1", "--paths", "8", "--dir", str(tmp_path)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 1 and "frg_create" in r.stderr
        assert os.listdir(tmp_path) == []
