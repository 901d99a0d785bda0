"""GPU tests of learned step control: the training kernels' gradient against the numpy oracle,
training progress, the classical fallback, and what the trained model gains on a held-out task
with the RK error test still accepting or rejecting every step."""
import sys

import numpy as np
import pytest

from finite_root_generation_amd import capi, tasks
from oracle import mlp_oracle as M
from tests.rootcmp import ROOT_TOL, abs_diff

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def records(tracker):
    fpm, g = tasks.seeded_task(10)
    say("[gpu] collect step log n=2048")
    tracker.set_opts(step_control=capi.STEP_CLASSIC, schedule=capi.SCHED_STATIC)
    rec = tracker.collect_step_log(fpm, g, seed=3, n=2048, cap=400_000)
    assert rec.shape[0] > 100_000 and np.all(np.isfinite(rec))
    return rec


def test_grad_kernel_matches_oracle(tracker, records):
    rng = np.random.default_rng(0)
    m = M.random_model(rng, records).astype(np.float32)
    sub = records[:3000]
    say("[gpu] step model gradient")
    g, loss = tracker.step_model_grad(m, sub)
    lo, go = M.loss_and_grad(m.astype(np.float64), sub)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)


def test_inlined_inference_matches_oracle(tracker, records):
    """The tracker's own inference code (FP16 weights, FP16 first-layer activations, FP32
    accumulation, as it proposes steps) against the float64 numpy forward pass, on 4,096 logged
    feature vectors and on a trained and a random model."""
    rng = np.random.default_rng(1)
    trained, _ = tracker.train_step_model(records, steps=300, batch=2048, lr=2e-3, seed=2)
    feats = records[rng.choice(records.shape[0], 4096, replace=False), :8]
    for name, m in (("trained", trained), ("random", M.random_model(rng, records).astype(np.float32))):
        tracker.set_step_model(m)
        y = tracker.step_model_eval(feats).astype(np.float64)
        yo = M.forward(m.astype(np.float64), feats)
        d = np.abs(y - yo) / m[M.YSD]  # in the model's normalised output units
        say(f"    {name} model: |y - oracle| / ysd median {np.median(d):.2e} max {d.max():.2e}; "
            f"spread of the oracle's y / ysd {np.std(yo) / m[M.YSD]:.2f}")
        assert np.all(np.isfinite(y))
        assert np.median(d) < 1e-3 and d.max() < 5e-3


def test_training_reduces_loss(tracker, records):
    say("[gpu] train 800 steps")
    model, loss = tracker.train_step_model(records, steps=800, batch=2048, lr=2e-3, seed=1)
    assert np.all(np.isfinite(model)) and np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.25 * loss[:10].mean()
    # the trained model beats predicting the mean target (normalised loss 0.5) on held-out data
    lo, _ = M.loss_and_grad(model.astype(np.float64), records[-20000:])
    assert lo < 0.1


def test_fallback_to_classical(tracker, records):
    """A model whose proposal overflows falls back to the classical rule: identical paths."""
    fpm, g = tasks.seeded_task(12)
    n = 512
    bad = M.random_model(np.random.default_rng(2), records).astype(np.float32)
    bad[M.YMU] = 1e6
    tracker.set_opts(step_control=capi.STEP_CLASSIC)
    a = tracker.track(fpm, g, seed=5, n=n)
    tracker.set_step_model(bad)
    tracker.set_opts(step_control=capi.STEP_LEARNED)
    try:
        say("[gpu] learned with overflowing model")
        b = tracker.track(fpm, g, seed=5, n=n)
    finally:
        tracker.set_opts(step_control=capi.STEP_CLASSIC)
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[0].view(np.float64), b[0].view(np.float64), equal_nan=True)


def test_learned_control_gain_on_held_out_task(tracker):
    """A model trained on ~3M step records of task 10 against the classical controller on 2^19
    paths of task 11 (persistent schedule, so the launch is not dominated by its slowest paths):
    fewer RK attempts and fewer solves (tangents + Newton iterations) per path, a shorter tracking
    kernel (best of two launches each), the same finite roots at 1e-8 for nearly every path."""
    fpm10, g10 = tasks.seeded_task(10)
    tracker.set_opts(step_control=capi.STEP_CLASSIC, schedule=capi.SCHED_PERSISTENT)
    say("[gpu] collect step log of task 10, n=32768")
    rec = tracker.collect_step_log(fpm10, g10, seed=3, n=32768, cap=4_000_000)
    model, _ = tracker.train_step_model(rec, steps=4000, batch=4096, lr=2e-3, seed=1)
    tracker.set_step_model(model)
    fpm, g = tasks.seeded_task(11)
    n = 1 << 19
    res, ms = {}, {}
    try:
        for name, sc in (("classical", capi.STEP_CLASSIC), ("learned", capi.STEP_LEARNED)):
            tracker.set_opts(step_control=sc, schedule=capi.SCHED_PERSISTENT)
            times = []
            for rep in range(2):
                say(f"[gpu] {name} n={n} run {rep}")
                res[name] = tracker.track(fpm, g, seed=4, n=n)
                times.append(res[name][3].kernel_ms)
            ms[name] = min(times)
    finally:
        tracker.set_opts(step_control=capi.STEP_CLASSIC, schedule=capi.SCHED_STATIC)
    (fa, ea, sa, _), (fb, eb, sb, _) = res["classical"], res["learned"]
    ma, mb = tasks.finite_mask(fa, ea), tasks.finite_mask(fb, eb)
    both = ma & mb
    recall = (abs_diff(fb[both], fa[both]) <= ROOT_TOL).sum() / ma.sum()
    rk_a, rk_b = sa["rk_attempts"].mean(), sb["rk_attempts"].mean()
    solves_a = (sa["tangent_solves"] + sa["newton_its"]).mean()
    solves_b = (sb["tangent_solves"] + sb["newton_its"]).mean()
    say(f"    {rec.shape[0]} records; classical rk/path {rk_a:.2f} solves/path {solves_a:.1f} kernel {ms['classical']:.0f} "
        f"ms; learned rk/path {rk_b:.2f} solves/path {solves_b:.1f} kernel {ms['learned']:.0f} ms; same root on "
        f"the same path {recall:.4f}; finite {ma.sum()} vs {mb.sum()}")
    assert recall >= 0.99
    assert abs(int(mb.sum()) - int(ma.sum())) <= 0.005 * n
    assert rk_b < rk_a
    assert solves_b < solves_a
    assert ms["learned"] < ms["classical"]
