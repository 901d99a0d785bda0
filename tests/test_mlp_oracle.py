"""Pins the numpy oracle of the step-size model (oracle/mlp_oracle.py) by finite differences,
and checks its parameter layout against the C ABI's constants."""
import numpy as np

from finite_root_generation_amd import capi
from oracle import mlp_oracle as M


def test_layout_matches_abi():
    assert M.NW == capi.STEP_MODEL_WEIGHTS == 4801
    assert M.SIZE == capi.STEP_MODEL_SIZE == 4819
    assert M.NIN + 1 == capi.STEP_RECORD


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    rec = np.concatenate([rng.normal(size=(40, 8)), rng.normal(size=(40, 1))], axis=1)
    m = M.random_model(rng, rec)
    loss, g = M.loss_and_grad(m, rec)
    h = 1e-6
    for p in rng.choice(M.NW, 60, replace=False).tolist() + [M.B3, M.W1, M.B2 - 1]:
        mp, mm = m.copy(), m.copy()
        mp[p] += h
        mm[p] -= h
        fd = (M.loss_and_grad(mp, rec)[0] - M.loss_and_grad(mm, rec)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-7 + 1e-5 * abs(fd), p


def test_forward_consistent_with_loss():
    rng = np.random.default_rng(1)
    rec = np.concatenate([rng.normal(size=(30, 8)), rng.normal(size=(30, 1))], axis=1)
    m = M.random_model(rng, rec)
    y = M.forward(m, rec[:, :8])
    t = rec[:, 8]
    loss = 0.5 * np.mean(((y - t) / m[M.YSD]) ** 2)
    assert abs(loss - M.loss_and_grad(m, rec)[0]) < 1e-12
