"""GPU parity of the 16-lane tracker (one path per 16 lanes, 4 paths per wave, FP64) against the
CPU oracle on identical inputs, as tests/test_gpu_parity.py does for the 8-lane one.

Stage level: one tangent solve (IVP, both t-step modes) and one Newton update through
frg_solve_step with lanes_per_path = 16, against the oracle's solves.  Path level: whole tracks
at both register budgets (1 and 2 waves per SIMD), codes and roots against the oracle to 1e-8
absolute, a disagreeing path accepted only where the oracle shows it to be rounding-sensitive.
Launch level: the schedule and the register budget do not change a path's arithmetic.
"""
import sys

import numpy as np
import pytest

from finite_root_generation_amd import capi, tasks
from oracle import oracle as O
from tests.rootcmp import ROOT_TOL, abs_diff, found_anywhere
from tests.test_gpu_parity import _compare_tracks, _stage_inputs, _starts, rel_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def t16():
    t = capi.Tracker(0, lanes_per_path=16)
    yield t
    t.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_solve16_tangent(t16, mode):
    n = 64
    Y, U, W, spm = _stage_inputs(n)
    say(f"[gpu] 16-lane solve_step tangent mode={mode} n={n}")
    x = t16.solve_step(mode, Y, U, W, spm, tasks.REF_FPM, tasks.REF_GAMMA)
    for i in range(n):
        ref = O.ivp(mode, Y[i], W[i], U[i], spm[i], tasks.REF_FPM, tasks.REF_GAMMA)
        assert rel_err(x[i], ref) < 1e-9, (i, rel_err(x[i], ref))


def test_solve16_newton(t16):
    n = 64
    Y, U, W, spm = _stage_inputs(n)
    tg = np.real(Y[:, 22]) + 0.01
    say(f"[gpu] 16-lane solve_step newton n={n}")
    x = t16.solve_step(2, Y, U, None, spm, tasks.REF_FPM, tasks.REF_GAMMA, tgoal=tg)
    for i in range(n):
        YN, res, its = O.newton(tg[i], W[i], U[i], spm[i], tasks.REF_FPM, tasks.REF_GAMMA, 1, 0.0, 1, Y[i])
        assert rel_err(x[i], YN - Y[i]) < 1e-9, (i, rel_err(x[i], YN - Y[i]))


def test_solve16_rejects_fp32(t16):
    Y, U, W, spm = _stage_inputs(4)
    with pytest.raises(capi.FrgError):
        t16.solve_step(0, Y, U, W, spm, tasks.REF_FPM, tasks.REF_GAMMA, precision=capi.PREC_MIXED)


@pytest.mark.parametrize("waves", [1, 2])
def test_track16_against_oracle(t16, waves):
    n = 512
    fpm, gamma = tasks.seeded_task(0)
    say(f"[gpu] 16-lane track, {waves} wave(s) per SIMD, n={n}")
    t16.set_opts(waves_per_simd=waves)
    try:
        fpt_g, err_g, st_g, summ = t16.track(fpm, gamma, seed=7, n=n)
    finally:
        t16.set_opts(waves_per_simd=0)
    assert summ.lanes_per_path == 16 and summ.waves_per_simd == waves
    oo = O.default_opts()
    fpt_o, err_o, st_o, _ = O.homotopy(oo, fpm, gamma, 7, 0, n)
    spt, spm = _starts(7, n)
    _compare_tracks(fpt_g, err_g, fpt_o, err_o, explain=(oo, fpm, gamma, spt, spm, 7, 0))


def test_track16_gaussian_learned_against_oracle():
    """The configuration the benchmark runs (Gaussian starts, persistent schedule, learned step
    proposals) with 16 lanes per path, the step log for the model collected by the 16-lane kernel
    itself.  The learned proposals change the steps, so the oracle runs with classical control and
    the end points are compared: nearly every path ends on the oracle's root, and nearly every
    oracle root is among the GPU's."""
    n = 512
    fpm, gamma = tasks.seeded_task(5)
    oo = O.default_opts(start_mode=1)
    with capi.Tracker(0, lanes_per_path=16, schedule=capi.SCHED_PERSISTENT, start_mode=capi.START_GAUSSIAN) as t:
        fpm10, g10 = tasks.seeded_task(10)
        say("[gpu] 16-lane step log of task 10, n=4096")
        rec = t.collect_step_log(fpm10, g10, seed=3, n=4096, cap=1_000_000)
        assert rec.shape[0] > 100_000 and np.all(np.isfinite(rec))
        model, _ = t.train_step_model(rec, steps=1500, batch=4096, lr=2e-3, seed=1)
        t.set_step_model(model)
        t.set_opts(step_control=capi.STEP_LEARNED)
        say(f"[gpu] 16-lane persistent learned track with Gaussian starts n={n}")
        fpt_g, err_g, st_g, _ = t.track(fpm, gamma, seed=13, n=n)
    fpt_o, err_o, st_o, _ = O.homotopy(oo, fpm, gamma, 13, 0, n)
    ok_g, ok_o = tasks.finite_mask(fpt_g, err_g), tasks.finite_mask(fpt_o, err_o)
    same = ok_g & ok_o & (abs_diff(fpt_g, fpt_o) <= ROOT_TOL)
    zg = fpt_g[ok_g]
    zg = zg[np.argsort(zg[:, 0].real)]
    anywhere = np.array([found_anywhere(zg, x) for x in fpt_o[ok_o]])
    say(f"    finite gpu {ok_g.sum()} oracle {ok_o.sum()}; same root on the same path {same.sum()}; oracle roots "
        f"found anywhere by the GPU {anywhere.sum()} of {ok_o.sum()}; rk attempts per path learned "
        f"{st_g['rk_attempts'].mean():.1f}, oracle classical {st_o['rk_attempts'].mean():.1f}")
    assert ok_g.sum() >= 0.97 * ok_o.sum()
    assert same.sum() >= 0.95 * ok_o.sum()
    assert anywhere.mean() >= 0.97


def test_track16_schedule_and_budget_do_not_change_results(t16):
    """A path's arithmetic does not depend on the schedule or on the register budget of the
    kernel: bit-identical results, statistics included."""
    n = 1000
    fpm, gamma = tasks.seeded_task(3)
    say(f"[gpu] 16-lane static vs persistent vs 2 waves per SIMD n={n}")
    a = t16.track(fpm, gamma, seed=9, n=n)
    t16.set_opts(schedule=capi.SCHED_PERSISTENT)
    try:
        b = t16.track(fpm, gamma, seed=9, n=n)
        t16.set_opts(waves_per_simd=2)
        c = t16.track(fpm, gamma, seed=9, n=n)
    finally:
        t16.set_opts(schedule=capi.SCHED_STATIC, waves_per_simd=0)
    for r in (b, c):
        assert np.array_equal(a[1], r[1])
        assert np.array_equal(a[0].view(np.float64), r[0].view(np.float64), equal_nan=True)
        assert np.array_equal(a[2], r[2])


def test_track16_unsupported_combinations():
    with pytest.raises(capi.FrgError):
        capi.Tracker(0, lanes_per_path=16, precision=capi.PREC_MIXED)
    with pytest.raises(capi.FrgError):
        capi.Tracker(0, lanes_per_path=8, waves_per_simd=2)
