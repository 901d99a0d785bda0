"""GPU tests of the batched launch: many task instances in one persistent launch replayed from a
cached hipGraph (each replay checked against an eager launch, with other GPU work between the
replays), and the autotuner's choice of launch shape."""
import sys

import numpy as np
import pytest

from finite_root_generation_amd import capi, tasks
from tests.rootcmp import ROOT_TOL, abs_diff

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _instances(counts, base=20):
    fpms, gams = zip(*[tasks.seeded_task(base + k) for k in range(len(counts))])
    seeds = [5 + k for k in range(len(counts))]
    pids = [1000 * k for k in range(len(counts))]
    return np.array(fpms), np.array(gams), seeds, pids, counts


def _same(a, b):
    return (np.array_equal(a[1], b[1]) and np.array_equal(a[0].view(np.float64), b[0].view(np.float64))
            and np.array_equal(a[2], b[2]))


def test_graph_replays_equal_eager(tracker):
    eager = capi.Tracker(0, schedule=capi.SCHED_PERSISTENT)
    graph = capi.Tracker(0, schedule=capi.SCHED_PERSISTENT, use_graph=1)
    try:
        for counts in ([300, 1000, 64], [17, 512]):
            inst = _instances(counts)
            say(f"[gpu] batch via graph, counts {counts}")
            ref = eager.track_batch(*inst)
            for rep in range(3):
                got = graph.track_batch(*inst)
                assert _same(got, ref), (counts, rep)
                # other GPU work between replays
                tracker.track(tasks.REF_FPM, tasks.REF_GAMMA, seed=rep, n=96)
    finally:
        eager.close()
        graph.close()


def test_autotune_keeps_the_fastest_shape():
    t = capi.Tracker(0)
    try:
        fpm, gamma = tasks.seeded_task(21)
        say("[gpu] autotune on 8192 paths")
        rows, chosen = t.autotune(fpm, gamma, seed=3, n_sample=8192, first_path_id=1 << 40)
        for r in rows:
            say(f"    sched {r['schedule']} block {r['block_size']} groups/CU {r['groups_per_cu']} "
                f"lanes {r['lanes_per_path']} waves/SIMD {r['waves_per_simd']}: "
                f"{r['kernel_ms']:.1f} ms, {r['paths_per_s']:.0f} paths/s")
        assert len(rows) >= 4
        assert {r["lanes_per_path"] for r in rows} == {8, 16}  # FP64: both path widths are candidates
        best = min(rows, key=lambda r: r["kernel_ms"])
        assert chosen == (best["schedule"], best["block_size"], best["groups_per_cu"], best["lanes_per_path"],
                          best["waves_per_simd"])
        # the chosen shape tracks the same paths as the default launch: bit for bit with the same
        # path width (the schedule does not change a path's arithmetic), to the root tolerance
        # with the other width (the 16-lane solve sums in another order)
        a = t.track(fpm, gamma, seed=9, n=256)
        with capi.Tracker(0) as d:
            b = d.track(fpm, gamma, seed=9, n=256)
        if chosen[3] == 8:
            assert np.array_equal(a[1], b[1])
            assert np.array_equal(a[0].view(np.float64), b[0].view(np.float64))
        else:
            both = tasks.finite_mask(*a) & tasks.finite_mask(*b)
            same = both & (abs_diff(a[0], b[0]) <= ROOT_TOL)
            say(f"    16 vs 8 lanes: {int(same.sum())} of {int(both.sum())} common finite roots agree")
            assert same.sum() >= 0.98 * both.sum() and (a[1] == b[1]).mean() >= 0.98
    finally:
        t.close()


def test_device_launches_on_two_streams():
    """Back-to-back frg_track_device calls on two caller streams, no synchronisation between
    them: the second call must not overwrite the instance buffers or the work-queue counter of
    the first launch while it runs, so both results equal separate synchronous launches."""
    import torch

    dev = torch.device("cuda:0")
    (fa, ga), (fb, gb) = tasks.seeded_task(31), tasks.seeded_task(32)
    n = 4096
    with capi.Tracker(0, schedule=capi.SCHED_PERSISTENT) as t:
        ref_a = t.track(fa, ga, seed=1, n=n, want_stats=False)
        ref_b = t.track(fb, gb, seed=2, n=n, want_stats=False)
        bufs = [(torch.empty((n, 21), dtype=torch.complex128, device=dev), torch.empty(n, dtype=torch.int32, device=dev))
                for _ in range(2)]
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        torch.cuda.synchronize()
        say(f"[gpu] frg_track_device on two streams, n={n} each")
        t.track_device(fa, ga, 1, n, bufs[0][0].data_ptr(), bufs[0][1].data_ptr(), stream=s1.cuda_stream)
        t.track_device(fb, gb, 2, n, bufs[1][0].data_ptr(), bufs[1][1].data_ptr(), stream=s2.cuda_stream)
        s1.synchronize()
        s2.synchronize()
        for (f, e), ref in zip(bufs, (ref_a, ref_b)):
            assert np.array_equal(e.cpu().numpy(), ref[1])
            assert np.array_equal(f.cpu().numpy().view(np.float64), ref[0].view(np.float64), equal_nan=True)


def test_batch_with_empty_and_ragged_instances(tracker):
    """Instances of 0, 1 and 13 paths between larger ones: every instance equals its own launch,
    and an empty instance takes no output rows."""
    counts = [37, 0, 1, 13, 0, 250]
    inst = _instances(counts, base=60)
    say(f"[gpu] batch with counts {counts}")
    with capi.Tracker(0, schedule=capi.SCHED_PERSISTENT) as t:
        fpt, err, st, summ = t.track_batch(*inst)
        assert fpt.shape[0] == sum(counts)
        off = 0
        for k, c in enumerate(counts):
            if c:
                one = t.track(inst[0][k], inst[1][k], seed=inst[2][k], n=c, first_path_id=inst[3][k])
                assert _same((fpt[off:off + c], err[off:off + c], st[off:off + c]), one), k
            off += c
