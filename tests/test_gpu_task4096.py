"""The 4,096-path workload (precision points of the seed-0 task, start seed 0) on every path
against the oracle: the plain FP64 tracker per path (codes, roots to 1e-8 absolute, finite-root
and failure counts, deduplicated root sets), and the mixed-precision tracker on the same inputs
(per-root error against the FP64 oracle, finite-root recall).  The start draws are the standard
complex Gaussian ones of this workload's definition, and the reference's uniform [-5, 5]^2 ones
(frgkinematics.cu:7-8) as a second case."""
import sys

import numpy as np
import pytest

from finite_root_generation_amd import capi, tasks
from oracle import oracle as O
from tests.rootcmp import ROOT_TOL, abs_diff, found_anywhere, rel_diff
from tests.test_gpu_parity import rounding_sensitive

pytestmark = pytest.mark.gpu
N = 4096
SEED = 0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[capi.START_GAUSSIAN, capi.START_UNIFORM], ids=["gaussian", "uniform"])
def oracle_run(request):
    mode = request.param
    fpm, gamma = tasks.seeded_task(0)
    say(f"[gpu] 4096-path workload (start mode {mode}): oracle on {N} paths")
    oo = O.default_opts(start_mode=mode)
    fpt, err, st, _ = O.homotopy(oo, fpm, gamma, SEED, 0, N)
    return fpm, gamma, oo, fpt, err, mode


def test_plain_tracker_every_path(oracle_run):
    fpm, gamma, oo, fpt_o, err_o, mode = oracle_run
    say(f"[gpu] 4096-path workload: plain FP64 tracker")
    with capi.Tracker(0, start_mode=mode) as t:
        fpt_g, err_g, _, _ = t.track(fpm, gamma, seed=SEED, n=N)
    fin_g, fin_o = tasks.finite_mask(fpt_g, err_g), tasks.finite_mask(fpt_o, err_o)
    both = fin_g & fin_o
    d = abs_diff(fpt_g, fpt_o)
    dr = rel_diff(fpt_g, fpt_o)
    agree = (err_g == err_o) & (~both | (d <= ROOT_TOL)) & (fin_g == fin_o)
    bad = np.nonzero(~agree)[0]
    explained = []
    for i in bad[:64]:
        spt, spm = O.generate_start(oo, SEED, int(i))
        if rounding_sensitive(oo, fpm, gamma, spt, spm, SEED, int(i)):
            explained.append(int(i))
    cg, co = np.bincount(err_g, minlength=6), np.bincount(err_o, minlength=6)
    say(f"    codes gpu {cg.tolist()} oracle {co.tolist()}; finite roots gpu {fin_g.sum()} oracle {fin_o.sum()}; "
        f"agree {agree.sum()}/{N}; disagreeing {bad.tolist()[:16]}, rounding-sensitive in the oracle {explained}; "
        f"max |dx| on agreeing paths {d[both & agree].max():.2e} (relative {dr[both & agree].max():.2e})")
    assert len(bad) <= 0.01 * N
    assert set(bad.tolist()) == set(explained), "GPU/oracle disagreement on a rounding-stable path"
    # counts: identical except where an explained path moved between codes (each such path moves
    # one count from one code to another)
    assert np.abs(cg - co).sum() <= 2 * len(bad)
    assert abs(int(fin_g.sum()) - int(fin_o.sum())) <= len(bad)
    # deduplicated root sets (the reference's rule, tolerance 1e-6)
    set_g, _ = O.sort_dedupe_merge(np.zeros((0, 21)), fpt_g, err_g, tol=1e-6)
    set_o, _ = O.sort_dedupe_merge(np.zeros((0, 21)), fpt_o, err_o, tol=1e-6)
    say(f"    distinct roots gpu {len(set_g)} oracle {len(set_o)}")
    assert abs(len(set_g) - len(set_o)) <= len(bad)
    set_g = set_g[np.argsort(set_g[:, 0].real)]
    lost = sum(not found_anywhere(set_g, r) for r in set_o)
    say(f"    oracle roots not in the GPU set at 1e-8: {lost}")
    assert lost <= len(bad)


def test_mixed_precision_against_fp64_oracle(oracle_run):
    """FP32 Jacobian and RK stages, FP64 corrector, end game and sharpening: per-root error of the
    mixed tracker against the FP64 oracle on paths that end on the same root, and the recall."""
    fpm, gamma, oo, fpt_o, err_o, mode = oracle_run
    say(f"[gpu] 4096-path workload: mixed precision tracker")
    with capi.Tracker(0, precision=capi.PREC_MIXED, start_mode=mode) as t:
        fpt_m, err_m, st_m, _ = t.track(fpm, gamma, seed=SEED, n=N)
    fin_m, fin_o = tasks.finite_mask(fpt_m, err_m), tasks.finite_mask(fpt_o, err_o)
    both = fin_m & fin_o
    d = abs_diff(fpt_m, fpt_o)
    same_root = both & (rel_diff(fpt_m, fpt_o) <= 1e-6)
    recall = (both & (d <= ROOT_TOL)).sum() / max(1, fin_o.sum())
    say(f"    finite roots mixed {fin_m.sum()} oracle {fin_o.sum()}; recall at 1e-8 {recall:.5f}; per-root error on "
        f"{same_root.sum()} shared roots: median {np.median(d[same_root]):.2e} max {d[same_root].max():.2e}; "
        f"fp32 solves {int(st_m['fp32_solves'].sum())} of {int((st_m['tangent_solves'] + st_m['newton_its']).sum())}")
    assert st_m["fp32_solves"].sum() > 0
    assert recall >= 0.98
    # shared roots are sharpened in FP64: their error against the oracle is far below the tolerance
    assert np.median(d[same_root]) < 1e-11
    assert np.mean(d[same_root] <= ROOT_TOL) >= 0.999
