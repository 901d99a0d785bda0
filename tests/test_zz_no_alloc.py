"""The allocation-free batch entry point (frg_track_batch_noalloc, device buffers, fixed per-GPU
scratch) against frg_track_batch through the C ABI: the same roots, codes and per-path counters
bit for bit, on two calls in a row whose second is smaller, for both schedules; and the limit on
the number of instances."""
import ctypes as ct
import sys

import numpy as np
import pytest
import torch

from finite_root_generation_amd import capi, tasks

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_noalloc(t, fpms, gams, seeds, pids, counts):
    dev = torch.device("cuda:0")
    k, n = len(counts), int(np.sum(counts))
    d_fpm = torch.from_numpy(np.ascontiguousarray(fpms, np.complex128)).to(dev)
    d_fpt = torch.empty((max(n, 1), 21), dtype=torch.complex128, device=dev)
    d_err = torch.empty((max(n, 1),), dtype=torch.int32, device=dev)
    d_st = torch.empty((max(n, 1), 8), dtype=torch.int32, device=dev)
    gam = np.ascontiguousarray(gams, np.complex128)
    sd = np.ascontiguousarray(seeds, np.uint64)
    pid = np.ascontiguousarray(pids, np.uint64)
    cnt = np.ascontiguousarray(counts, np.int64)
    torch.cuda.synchronize()
    rc = capi.lib().frg_track_batch_noalloc(t._h, k, ct.c_void_p(d_fpm.data_ptr()), capi._p(gam), capi._p(sd),
                                            capi._p(pid), capi._p(cnt), ct.c_void_p(d_fpt.data_ptr()),
                                            ct.c_void_p(d_err.data_ptr()), ct.c_void_p(d_st.data_ptr()), None)
    assert rc == 0, capi.lib().frg_strerror(rc)
    torch.cuda.synchronize()
    return d_fpt[:n].cpu().numpy(), d_err[:n].cpu().numpy(), d_st[:n].cpu().numpy()


@pytest.mark.parametrize("sched", [capi.SCHED_STATIC, capi.SCHED_PERSISTENT], ids=["static", "persistent"])
def test_noalloc_batch_equals_batch(sched):
    with capi.Tracker(0, schedule=sched, start_mode=capi.START_GAUSSIAN) as t:
        for call, counts in enumerate(([700, 1, 0, 2048, 333], [300, 90])):
            k = len(counts)
            fpms, gams = zip(*[tasks.seeded_task(40 + 10 * call + i) for i in range(k)])
            fpms, gams = np.array(fpms), np.array(gams)
            seeds = 7 + np.arange(k)
            pids = 1000 * np.arange(k) + call
            say(f"[gpu] no-alloc batch, schedule {sched}, call {call}: {k} instances, {sum(counts)} paths")
            f_ref, e_ref, s_ref, _ = t.track_batch(fpms, gams, seeds, pids, counts)
            f_got, e_got, s_got = run_noalloc(t, fpms, gams, seeds, pids, counts)
            assert np.array_equal(e_ref, e_got)
            assert np.array_equal(f_ref.view(np.float64), f_got.view(np.float64), equal_nan=True)
            assert np.array_equal(s_ref.view(np.int32).reshape(-1, 8), s_got)


def test_noalloc_batch_limit():
    with capi.Tracker(0) as t:
        k = 65
        fpm = torch.zeros((k, 30), dtype=torch.complex128, device="cuda:0")
        gam = np.ones(k, np.complex128)
        sd = np.zeros(k, np.uint64)
        cnt = np.ones(k, np.int64)
        rc = capi.lib().frg_track_batch_noalloc(t._h, k, ct.c_void_p(fpm.data_ptr()), capi._p(gam), capi._p(sd),
                                                capi._p(sd), capi._p(cnt), None, None, None, None)
        assert rc == -5 and b"limit" in capi.lib().frg_strerror(rc)
