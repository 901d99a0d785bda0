"""GPU tests of the round driver (finite_root_generation_amd/frg_rounds): the reference's resume
files fptColl / errorCodesAgr / progress (frg.cu:295-336, 467-512), a resumed run, the same
collection recomputed through the C ABI from the same path ids, and a run with the learned
step-size model trained by the driver itself."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from finite_root_generation_amd import capi, tasks
from tests import stephenson_np as S
from tests.rootcmp import ROOT_TOL, abs_diff

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(capi.LIB_PATH), "frg_rounds")
N = 2048


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_rounds(d, rounds, *extra):
    say(f"[gpu] frg_rounds {rounds} --paths {N} {' '.join(extra)}")
    r = subprocess.run([EXE, str(rounds), "--paths", str(N), "--dir", str(d), *extra], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def expected(tracker, rounds, seed=1):
    """The collection and codes recomputed from the same path ids (round k: ids k*N ...)."""
    coll = np.zeros((0, 21), np.complex128)
    codes = []
    for k in range(rounds):
        fpt, err, _, _ = tracker.track(tasks.REF_FPM, tasks.REF_GAMMA, seed=seed, n=N, first_path_id=k * N)
        coll, err = capi.sort_dedupe_merge(coll, fpt, err, 1e-6)
        codes.append(err)
    return coll, np.concatenate(codes)


def read_files(d):
    coll = np.fromfile(os.path.join(d, "fptColl"), dtype=np.complex128).reshape(-1, 21)
    codes = np.fromfile(os.path.join(d, "errorCodesAgr"), dtype=np.int32)
    progress = open(os.path.join(d, "progress")).read()
    return coll, codes, progress


def test_rounds_files_and_resume(tracker, tmp_path):
    assert os.path.exists(EXE), "frg_rounds is not built"
    run_rounds(tmp_path, 2, "--persistent")
    coll, codes, progress = read_files(tmp_path)
    assert codes.shape == (2 * N,)
    assert os.path.getsize(os.path.join(tmp_path, "fptColl")) == coll.shape[0] * 336
    ecoll, ecodes = expected(tracker, 2)
    assert np.array_equal(codes, ecodes)
    assert np.array_equal(coll.view(np.float64), ecoll.view(np.float64))
    # the collection is sorted on Re z0 and holds one copy of every code-0 root
    assert np.all(np.diff(coll[:, 0].real) >= 0)
    assert coll.shape[0] == int((codes == 0).sum())
    assert "-----Progress statistics for round 1-----" in progress
    assert "-----Progress statistics for round 2-----" in progress

    # resume: one more round continues with the next path ids
    out = run_rounds(tmp_path, 1)
    assert "Previous rounds: 2" in out and "Round 3" in out
    coll, codes, progress = read_files(tmp_path)
    ecoll, ecodes = expected(tracker, 3)
    assert np.array_equal(codes, ecodes)
    assert np.array_equal(coll.view(np.float64), ecoll.view(np.float64))
    blocks = progress.split("-----Progress statistics for round ")
    assert len(blocks) == 4 and blocks[3].startswith("3-----")
    last = blocks[3]
    roots = int(re.search(r"---In Total---\nRoots collected:\s+(\d+)", last).group(1))
    dups = int(re.search(r"---In Total---\n.*\nDuplicates:\s+(\d+)", last).group(1))
    trials = int(re.search(r"---In Total---\n.*\n.*\nTrials:\s+(\d+)", last).group(1))
    assert roots == int((codes == 0).sum()) and dups == int((codes == 6).sum()) and trials == 3 * N
    est = float(re.search(r"Roots collected \(est\.\):\s+([0-9.eE+-]+)%", last).group(1))
    assert abs(est - 100 * capi.completeness(roots, dups)) <= 1e-3 * max(1.0, est)


def test_rounds_with_learned_step_control(tracker, tmp_path):
    """--learned: the driver trains the model on its own step logs (path ids 2^62 on), then tracks
    the round with it.  The collection holds roots of the target, and nearly every root of the
    plain run of the same path ids."""
    out = run_rounds(tmp_path, 1, "--persistent", "--learned", "--train-paths", "4096")
    assert "Step model trained on" in out
    coll, codes, progress = read_files(tmp_path)
    assert codes.shape == (N,)
    assert coll.shape[0] == int((codes == 0).sum())
    assert np.all(np.diff(coll[:, 0].real) >= 0)
    F = S.fsys(np.concatenate([coll, np.ones((coll.shape[0], 1))], axis=1), tasks.REF_FPM[None, :])
    rel = np.max(np.abs(F), axis=1) / (1.0 + np.max(np.abs(coll), axis=1)) ** 3
    assert rel.max() < 1e-9
    ecoll, _ = expected(tracker, 1)
    found = 0
    for r in ecoll:
        found += bool(abs_diff(coll, r[None]).min() <= ROOT_TOL)
    say(f"    learned collection {coll.shape[0]} roots, plain {ecoll.shape[0]}; plain roots found {found}")
    assert found >= 0.97 * ecoll.shape[0]


def test_rounds_per_launch_gives_the_same_files(tracker, tmp_path):
    """--rounds-per-launch 2 over three rounds (one launch of two rounds, one of one): the files
    equal those of one launch per round, path id for path id."""
    run_rounds(tmp_path, 3, "--persistent", "--rounds-per-launch", "2")
    coll, codes, progress = read_files(tmp_path)
    ecoll, ecodes = expected(tracker, 3)
    assert np.array_equal(codes, ecodes)
    assert np.array_equal(coll.view(np.float64), ecoll.view(np.float64))
    assert progress.count("-----Progress statistics for round ") == 3


def test_resume_with_another_path_count_uses_fresh_ids(tracker, tmp_path):
    """A resumed run with a different --paths continues the path ids after the trials already
    recorded in errorCodesAgr, so no path is tracked twice: the files equal those of rounds over
    consecutive ids (N, then two rounds of M)."""
    m = N // 2 + 100
    run_rounds(tmp_path, 1)
    run_rounds(tmp_path, 2, "--paths", str(m))
    coll, codes, progress = read_files(tmp_path)
    assert codes.shape == (N + 2 * m,)
    ecoll = np.zeros((0, 21), np.complex128)
    ecodes = []
    for first, n in ((0, N), (N, m), (N + m, m)):
        fpt, err, _, _ = tracker.track(tasks.REF_FPM, tasks.REF_GAMMA, seed=1, n=n, first_path_id=first)
        ecoll, err = capi.sort_dedupe_merge(ecoll, fpt, err, 1e-6)
        ecodes.append(err)
    assert np.array_equal(codes, np.concatenate(ecodes))
    assert np.array_equal(coll.view(np.float64), ecoll.view(np.float64))
    assert progress.count("-----Progress statistics for round ") == 3


def test_failed_collection_write_keeps_the_previous_files(tmp_path):
    """fptColl is written to fptColl.tmp and renamed over the old one: when that write fails (here
    a directory of that name stands in the way), the run stops with an error and the files of the
    earlier rounds stay as they were."""
    run_rounds(tmp_path, 1)
    coll0, codes0, _ = read_files(tmp_path)
    os.mkdir(os.path.join(tmp_path, "fptColl.tmp"))
    say("[gpu] frg_rounds 1 with fptColl.tmp blocked")
    r = subprocess.run([EXE, "1", "--paths", str(N), "--dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 1 and "cannot write fptColl" in r.stderr, r.stdout + r.stderr
    coll1, codes1, _ = read_files(tmp_path)
    assert np.array_equal(coll0, coll1) and np.array_equal(codes0, codes1)
