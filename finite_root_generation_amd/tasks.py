"""Task (target parameter) definitions for the Stephenson II finite-root-generation workload.

Parameter layout (htpykinematics.cu:29-33): [P0..P7, Pc0..Pc7, Q1..Q7, Qc1..Qc7].

* ``REF_FPM`` / ``REF_GAMMA``: the reference's hard-coded generic target (frg.cu:149-178) and
  gamma (frg.cu:41).
* ``seeded_task(seed, real=False)``: a generic task (P, Pc independent in the unit disk, Q, Qc
  complex Gaussian) or, with real=True, a physical one (Pc = conj P, |Q| = 1, Qc = conj Q); gamma
  on the unit circle; all from numpy's PCG64 generator seeded with ``seed``.
* ``finite_mask``: which outputs count as finite roots.
"""
from __future__ import annotations

import numpy as np

REF_GAMMA = complex(0.39231132891081666, 1.0062265257923928)

REF_FPM = np.array([
    complex(4.241713408139979, -4.925334736425448), complex(1.448005491494222, -2.686249989236069),
    complex(2.287911194543938, -0.9123108488455074), complex(-1.870502483790705, -0.5778077414588303),
    complex(-0.7625329418400497, 0.9043788195368165), complex(-4.142822753796651, 1.697671322829592),
    complex(1.023004534182935, -3.026210035710708), complex(-2.522295755933046, 3.942421532145637),
    complex(-3.769870484797034, 0.7972377857916531), complex(-4.285631161563732, -0.5650246890525263),
    complex(-2.166030800874488, 0.8836055774820508), complex(4.579260949000442, -2.074013737196861),
    complex(4.801289733755713, -1.460721841824302), complex(2.041266673362720, 2.962149232597863),
    complex(3.989999692714399, -0.2328295911150420), complex(-3.996448356004510, 4.426854856711802),
    complex(2.315936719955769, -0.5527803857724418), complex(-3.018911857947673, 0.9692384144638879),
    complex(-0.4906347357049512, 2.994316903874402), complex(0.2397118512793472, -1.341440197750693),
    complex(4.942917423339324, -2.415216683074435), complex(1.599536873563615, 2.367319121024760),
    complex(4.229853181495468, -1.149104287738972), complex(-0.2180298061101595, 3.585004213517870),
    complex(4.311063203842579, -0.03375595047392820), complex(4.938148066901551, 4.551701974206413),
    complex(-4.271214442569748, 0.1355349863802999), complex(-0.7377434797048270, 1.831738837919296),
    complex(0.8894614668331116, 0.8657473695398501), complex(4.036757863569754, 1.461164684969592),
], dtype=np.complex128)


def _disk(rng, n):
    return np.sqrt(rng.random(n)) * np.exp(2j * np.pi * rng.random(n))


def seeded_task(seed: int, real: bool = False):
    """Returns (fpm[30], gamma) of a seeded task.

    Generic (default): the precision points P0..P7 and their isotropic partners Pc0..Pc7 are
    drawn independently, uniform in the unit disk, and Q_j, Qc_j are standard complex Gaussian,
    so the target is a generic complex parameter point like the reference's own fpm
    (frg.cu:149-178, where Pc != conj P and Qc != 1/Q).
    real=True: a physical task, Pc = conj(P), Q_j on the unit circle and Qc = conj(Q); this is a
    special parameter point at which many paths end at singular roots.
    gamma is on the unit circle in both cases.
    """
    rng = np.random.default_rng(seed)
    P = _disk(rng, 8)
    if real:
        Q = np.exp(2j * np.pi * rng.random(7))
        Pc, Qc = np.conj(P), np.conj(Q)
    else:
        Pc = _disk(rng, 8)
        Q = (rng.normal(size=7) + 1j * rng.normal(size=7)) / np.sqrt(2)
        Qc = (rng.normal(size=7) + 1j * rng.normal(size=7)) / np.sqrt(2)
    gamma = complex(np.exp(2j * np.pi * rng.random()))
    fpm = np.concatenate([P, Pc, Q, Qc]).astype(np.complex128)
    return fpm, gamma


def finite_mask(fpt, err, big: float = 1e8):
    """Finite roots: code 0 and every component finite and below `big` in modulus."""
    fpt = np.asarray(fpt)
    ok = (np.asarray(err) == 0) & np.all(np.isfinite(fpt.view(np.float64).reshape(fpt.shape[0], -1)), axis=1)
    with np.errstate(invalid="ignore"):
        ok &= np.max(np.abs(fpt), axis=1) < big
    return ok
