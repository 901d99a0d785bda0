"""The device sort / dedupe / merge against the host frg_sort_dedupe_merge and the oracle's
restatement of SortingDeletingMerging (both pinned by the brute-force CPU test): bit-identical
collections and codes on adversarial sets -- clusters of near duplicates around tol, equal keys
across old and new roots, -0/+0 keys, NaN keys, failed paths -- on two rounds of real tracker
output, and on a round whose failed paths carry thousands of non-finite keys."""
import sys
import time

import numpy as np
import pytest
import torch

from finite_root_generation_amd import capi, tasks
from oracle import oracle as O

pytestmark = pytest.mark.gpu
NV = 21


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def device_merge(tracker, coll, fpt, err, tol=1e-6):
    dev = torch.device("cuda:0")
    m, n = coll.shape[0], fpt.shape[0]
    dc = torch.from_numpy(np.ascontiguousarray(coll, np.complex128).reshape(m, NV)).to(dev)
    df = torch.from_numpy(np.ascontiguousarray(fpt, np.complex128).reshape(n, NV)).to(dev)
    de = torch.from_numpy(np.ascontiguousarray(err, np.int32)).to(dev)
    out = torch.empty((max(m + n, 1), NV), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    k = tracker.sort_dedupe_merge_device(dc.data_ptr() if m else 0, m, df.data_ptr() if n else 0,
                                         de.data_ptr() if n else 0, n, out.data_ptr(), tol)
    torch.cuda.synchronize()
    return out[:k].cpu().numpy(), de.cpu().numpy()


def _sample(rng, base, k):
    pick = base[rng.integers(0, len(base), k)].copy()
    noise = rng.choice([0.0, 1e-8, 5e-7, 2e-6, 1.0], size=(k, 1))
    pick = pick + noise * (rng.normal(size=pick.shape) + 1j * rng.normal(size=pick.shape))
    z = rng.random(k)
    pick[z < 0.1, 0] = complex(0.0, 1.0)
    pick[(z >= 0.1) & (z < 0.2), 0] = complex(-0.0, 1.0)
    pick[(z >= 0.2) & (z < 0.25), 0] = complex(np.nan, 0.0)
    return pick


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.float64), b.view(np.float64), equal_nan=True)


def test_device_dedupe_matches_host_on_adversarial_sets(tracker):
    rng = np.random.default_rng(0)
    say("[gpu] device dedupe, 200 adversarial sets")
    for trial in range(200):
        base = rng.normal(size=(12, NV)) + 1j * rng.normal(size=(12, NV))
        m0, n = int(rng.integers(0, 30)), int(rng.integers(0, 40))
        old = _sample(rng, base, m0)
        old_err = np.where(rng.random(m0) < 0.1, 3, 0).astype(np.int32)
        coll, _ = capi.sort_dedupe_merge(np.zeros((0, NV), complex), old, old_err)
        fpt = _sample(rng, base, n)
        err = np.where(rng.random(n) < 0.15, rng.integers(1, 6, n), 0).astype(np.int32)
        if trial % 2 == 0:  # no code-0 NaN key: the per-run path
            err[np.isnan(fpt[:, 0].real) & (err == 0)] = 2
        h_coll, h_err = capi.sort_dedupe_merge(coll, fpt, err)
        d_coll, d_err = device_merge(tracker, coll, fpt, err)
        o_coll, o_err = O.sort_dedupe_merge(coll, fpt, err)
        assert np.array_equal(h_err, d_err), trial
        assert _same(h_coll, d_coll), trial
        assert np.array_equal(o_err, d_err), trial
        assert _same(o_coll, d_coll), trial


def test_device_dedupe_matches_host_on_tracker_rounds(tracker):
    fpm, gamma = tasks.seeded_task(5)
    n = 16384
    coll_h = coll_d = np.zeros((0, NV), complex)
    for r in range(2):
        say(f"[gpu] device dedupe, tracker round {r} n={n}")
        fpt, err, _, _ = tracker.track(fpm, gamma, seed=3, n=n, first_path_id=r * n)
        coll_o, err_o = O.sort_dedupe_merge(coll_h, fpt, err)
        coll_h, err_h = capi.sort_dedupe_merge(coll_h, fpt, err)
        coll_d, err_d = device_merge(tracker, coll_d, fpt, err)
        assert np.array_equal(err_h, err_d)
        assert _same(coll_h, coll_d)
        assert np.array_equal(err_o, err_d)
        assert _same(coll_o, coll_d)
        say(f"    collection {coll_h.shape[0]} roots, duplicates this round {(err_h == 6).sum()}")
    assert (err_h == 6).sum() > 0


def test_device_dedupe_empty(tracker):
    c, e = device_merge(tracker, np.zeros((0, NV), complex), np.zeros((0, NV), complex), np.zeros(0, np.int32))
    assert c.shape == (0, NV) and e.shape == (0,)


def test_device_dedupe_many_nonfinite_failed_paths(tracker):
    """Failed paths with NaN or infinite keys never break a run, so they form one run at the end of
    the merged order; the run is walked in index order once (no rescans).  5,000 of them among
    20,000 roots: same result as the host rule, in bounded time."""
    rng = np.random.default_rng(3)
    n, bad = 20000, 5000
    fpt = rng.normal(size=(n, NV)) + 1j * rng.normal(size=(n, NV))
    err = np.zeros(n, np.int32)
    idx = rng.choice(n, bad, replace=False)
    fpt[idx[: bad // 2], 0] = complex(np.nan, 0.0)
    fpt[idx[bad // 2:], 0] = complex(np.inf, 0.0)
    err[idx] = rng.integers(1, 6, bad)
    coll, _ = capi.sort_dedupe_merge(np.zeros((0, NV), complex), fpt[:5000], err[:5000])
    say("[gpu] device dedupe with 5,000 non-finite failed paths")
    device_merge(tracker, coll, fpt, err)  # warm-up (kernel loading)
    t0 = time.perf_counter()
    d_coll, d_err = device_merge(tracker, coll, fpt, err)
    dt = time.perf_counter() - t0
    h_coll, h_err = capi.sort_dedupe_merge(coll, fpt, err)
    say(f"    {dt * 1e3:.1f} ms")
    assert np.array_equal(h_err, d_err)
    assert _same(h_coll, d_coll)
    assert dt < 0.5
