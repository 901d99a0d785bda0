"""Benchmark: tracked paths/s on one MI355X, chosen tracker vs plain FP64, at equal finite-root recall.

Workload (one "step" = one pass over one batch): the Stephenson II finite-root-generation round
at full size, 1,048,576 paths; target task from seed 1 (finite_root_generation_amd.tasks
.seeded_task), start systems drawn by the tracker's Philox stream from seed 2 with standard
complex Gaussian free values.  Inputs are the 30 target parameters and the seed; outputs (roots,
codes, per-path counters) stay in HBM, so the timed region is the launch sequence alone
(tracking kernel, then the refinement + cognate kernel).  The same batch is run by
  * plain:  FP64, classical step control, static grid with one path per group, 8 lanes per path
            (the reference's launch: one fixed grid, every path tracked by its own thread group);
  * chosen: FP64, persistent grid with a device work queue, learned step-size control (model
            trained before the timed region on step logs of the same two tasks with other start
            seeds: the 4,096-path task 0 and this task 1), launch shape picked by the autotuner
            on a 2^19-path sample;
  * ablations (one launch each): persistent + classical, mixed FP32/FP64 + persistent + classical,
    and mixed + persistent + learned (all three changes together on one instance).
value = chosen paths/s (wall clock around the timed steps).  One JSON line on stdout.

The line also carries "batched", the many-instance workload: 64 task instances (seeded tasks
200..263, start seeds 300..363) with path counts log-uniform in [1,024, 262,144] (numpy seed 4),
FP32/FP64 mixed precision and learned step control, launch shape picked by the autotuner, all
instances in one persistent launch replayed from a hipGraph; beside it ("fp64_learned") the same
launch in FP64 with the chosen tracker's tuned shape, since FP32 solves do not run faster on this
GPU; and the plain way, each instance its own static FP64 launch with classical control, as the
reference runs one target at a time.

Parity and recall in the line: the CPU baseline's oracle run (the first paths of the same
workload) is compared path by path with the plain and chosen GPU runs (codes, roots to 1e-8
absolute), and the distinct finite roots of both GPU runs after the reference's 1e-6 dedupe
(frgsupport.cu:7-171) are compared as sets.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--paths N] [--cpu-sample N]
  python bench.py --traffic-probe     (for rocprofv3 --pmc runs: one chosen launch, then a plain
                                       copy of 1 GiB whose counter readings calibrate the counters)

roofline.achieved uses the frozen per-unit FLOP counts of finite_root_generation_amd/
flop_constants.json (this implementation's formulation; the reference formulation's count is
reported beside it); roofline.traffic comes from profiles/track_kernel_traffic.json when that file
is present: HBM bytes per chosen launch measured by two rocprofv3 --pmc runs (FETCH_SIZE,
WRITE_SIZE) of the traffic probe, each scaled by what the counter reads for the 1 GiB copy.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tracked paths/s on 1 MI355X, chosen vs plain FP64, at equal finite-root recall"
P64_TFLOPS = 78.6    # MI355X FP64 vector peak
P32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
ROOT_TOL = 1e-8      # absolute, max |dx| over the 21 components
DEDUPE_TOL = 1e-6    # the reference's duplicate tolerance (frgsupport.cu:7)
TUNE_SAMPLE = 1 << 19


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def flop_constants():
    with open(os.path.join(ROOT, "finite_root_generation_amd", "flop_constants.json")) as f:
        return json.load(f)


def run_variant(capi, torch, name, opts_kw, fpm, gamma, n, steps, warmup, seed, model=None, tune=None):
    """Runs `steps` timed launches after `warmup` untimed ones; returns outputs and timings.
    tune: (seed, n_sample) to pick the launch shape with the autotuner first."""
    t = capi.Tracker(0, **opts_kw)
    if model is not None:
        t.set_step_model(model)
    tuned = None
    if tune is not None:
        log(f"[bench] {name}: autotune on {tune[1]} paths")
        rows, chosen = t.autotune(fpm, gamma, seed=tune[0], n_sample=tune[1], first_path_id=1 << 40)
        tuned = {"rows": rows, "chosen": chosen}
        log(f"[bench] {name}: chosen {chosen}")
    dev = torch.device("cuda:0")
    fpt = torch.empty((n, 21), dtype=torch.complex128, device=dev)
    err = torch.empty((n,), dtype=torch.int32, device=dev)
    st = torch.empty((n, 8), dtype=torch.int32, device=dev)
    for w in range(warmup):
        log(f"[bench] {name}: warmup {w + 1}/{warmup}")
        t.track_device(fpm, gamma, seed, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        torch.cuda.synchronize()
    kern, refine = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        t.track_device(fpm, gamma, seed, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        kern.append(t.last_kernel_ms())
        refine.append(t.last_refine_ms())
        log(f"[bench] {name}: step {k + 1}/{steps} tracking kernel {kern[-1]:.1f} ms, refinement {refine[-1]:.1f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the reference's window ends with the results on the host (frg.cu:356-426): time the copy
    c0 = time.perf_counter()
    out = {"fpt": fpt.cpu().numpy(), "err": err.cpu().numpy()}
    copy_s = time.perf_counter() - c0
    out.update({
        "stats": st.cpu().numpy().view(capi.PATH_STATS_DTYPE).reshape(n),
        "ms_per_step": 1e3 * wall / steps, "kernel_ms": float(np.mean(kern)), "refine_ms": float(np.mean(refine)),
        "d2h_ms": 1e3 * copy_s, "tuned": tuned, "opts": t.opts,
    })
    t.close()
    return out


def path_flops(stats, unit):
    """(FP64 flops, FP32 flops) of the tracked paths from their counters and per-unit counts; the
    solves done in FP32 are counted separately."""
    tan = stats["tangent_solves"].astype(np.float64)
    nit = stats["newton_its"].astype(np.float64)
    solves = tan + nit
    f32_share = np.where(solves > 0, stats["fp32_solves"] / np.maximum(solves, 1), 0.0)
    per_solve = tan * unit["tangent_solve"] + nit * unit["newton_update"]
    f32 = float((per_solve * f32_share).sum())
    f64 = float((per_solve * (1 - f32_share)).sum() + stats["rk_attempts"].astype(np.float64).sum()
                * unit["rk_combination"])
    return f64, f32


def roofline(res, consts, hbm=None):
    sec = res["kernel_ms"] * 1e-3
    f64, f32 = path_flops(res["stats"], consts["compact"])
    r64, r32 = path_flops(res["stats"], consts["reference"])
    tmin = f64 / (P64_TFLOPS * 1e12) + f32 / (P32_TFLOPS * 1e12)
    achieved = (f64 + f32) / sec / 1e12
    peak = (f64 + f32) / tmin / 1e12 if tmin > 0 else P64_TFLOPS
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": hbm["bytes"] if hbm else None,
            "bound_detail": "FP64 vector ALU (the tracker issues no MFMA instruction); peak = MI355X FP64 vector "
                            "78.6 TFLOP/s (FP32 157.3 for FP32 solves)",
            "flops_source": "frozen per-unit counts, finite_root_generation_amd/flop_constants.json (compact "
                            "formulation) x the paths' own counters; tracking kernel time from HIP events",
            "flops_per_launch": {"fp64": f64, "fp32": f32},
            "reference_formulation": {"flops_per_launch": r64 + r32,
                                      "equivalent_tflops": round((r64 + r32) / sec / 1e12, 3)},
            "traffic_detail": hbm}


def distinct_roots(capi, res):
    """Distinct finite roots under the reference's rule (sort on Re z0, duplicates within 1e-6 on
    all 42 reals, frgsupport.cu:7-171), as the host merge into an empty collection."""
    t0 = time.perf_counter()
    coll, _ = capi.sort_dedupe_merge(np.zeros((0, 21)), res["fpt"], res["err"].copy(), tol=DEDUPE_TOL)
    from finite_root_generation_amd import tasks
    ok = tasks.finite_mask(coll, np.zeros(len(coll), np.int32))
    return coll[ok], time.perf_counter() - t0


def set_found(ref_set, other_set):
    """Number of ref_set's roots with a root of other_set within ROOT_TOL (absolute, every
    component); other_set sorted on Re z0.  Candidates are the other roots whose Re z0 lies within
    ROOT_TOL, visited by offset from the first one."""
    if len(ref_set) == 0 or len(other_set) == 0:
        return 0
    keys = other_set[:, 0].real
    lo = np.searchsorted(keys, ref_set[:, 0].real - ROOT_TOL, side="left")
    hi = np.searchsorted(keys, ref_set[:, 0].real + ROOT_TOL, side="right")
    found = np.zeros(len(ref_set), bool)
    for j in range(int((hi - lo).max(initial=0))):
        idx = lo + j
        live = (idx < hi) & ~found
        if not live.any():
            break
        cand = other_set[np.minimum(idx, len(other_set) - 1)]
        found |= live & (np.max(np.abs(cand - ref_set), axis=1) <= ROOT_TOL)
    return int(found.sum())


def set_recall(ref_set, other_set):
    """Fraction of ref_set's roots found in other_set (see set_found)."""
    return set_found(ref_set, other_set) / len(ref_set) if len(ref_set) else 1.0


def instance_sets(capi, tasks, fpt, err, offs):
    """Distinct finite roots of each instance under the reference's dedupe, sorted on Re z0."""
    out = []
    for k in range(len(offs) - 1):
        coll, _ = capi.sort_dedupe_merge(np.zeros((0, 21)), fpt[offs[k]:offs[k + 1]],
                                         err[offs[k]:offs[k + 1]].copy(), tol=DEDUPE_TOL)
        out.append(coll[tasks.finite_mask(coll, np.zeros(len(coll), np.int32))])
    return out


def sets_recall(ref_sets, other_sets):
    """Pooled recall over instances: roots of ref_sets found in the same instance's other set."""
    tot = sum(len(r) for r in ref_sets)
    return sum(set_found(r, o) for r, o in zip(ref_sets, other_sets)) / tot if tot else 1.0


def per_path_vs_oracle(tasks, g_fpt, g_err, o_fpt, o_err):
    """Same path ids: code histograms, paths whose codes agree and whose roots (when finite in
    both) agree to ROOT_TOL absolute."""
    n = len(o_err)
    g_fpt, g_err = g_fpt[:n], g_err[:n]
    both = tasks.finite_mask(g_fpt, g_err) & tasks.finite_mask(o_fpt, o_err)
    d = np.max(np.abs(g_fpt - o_fpt), axis=1)
    agree = (g_err == o_err) & (~both | (d <= ROOT_TOL))
    return {"paths": n, "agree": int(agree.sum()),
            "codes_gpu": np.bincount(g_err, minlength=6).tolist(), "codes_oracle": np.bincount(o_err, minlength=6).tolist(),
            "finite_gpu": int(tasks.finite_mask(g_fpt, g_err).sum()), "finite_oracle": int(tasks.finite_mask(o_fpt, o_err).sum()),
            "max_abs_dx_on_agreeing": float(d[both & agree].max()) if (both & agree).any() else 0.0,
            "disagreeing_ids": np.nonzero(~agree)[0][:32].tolist()}


def train_step_model(capi, tasks):
    """Step logs that the tracker writes on the workloads' own tasks (task 0 of the 4,096-path
    workload and task 1 of this one) with start seeds other than the measured one; one model
    trained on the GPU."""
    t0 = time.perf_counter()
    recs = []
    with capi.Tracker(0, start_mode=capi.START_GAUSSIAN) as t:
        for k, (task, s, n) in enumerate(((0, 100, 65536), (1, 101, 65536), (1, 102, 65536))):
            fpm, g = tasks.seeded_task(task)
            recs.append(t.collect_step_log(fpm, g, seed=s, n=n, cap=4_000_000))
        rec = np.concatenate(recs)[:10_000_000]
        model, loss = t.train_step_model(rec, steps=4000, batch=4096, lr=2e-3, seed=1)
    return model, {"records": int(rec.shape[0]), "train_steps": 4000, "final_loss": float(loss[-50:].mean()),
                   "logs": "tasks 0 and 1 (start seeds 100-102), 65,536 paths each",
                   "seconds": round(time.perf_counter() - t0, 2)}


def traffic(chosen_ms):
    """HBM bytes per chosen launch from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "track_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"bytes": t["hbm_bytes_per_launch"], "fetch_bytes": t["fetch_bytes"], "write_bytes": t["write_bytes"],
            "measured_on": t["kernel"], "source": "profiles/track_kernel_traffic.json"}


def chosen_opts(capi):
    return dict(precision=capi.PREC_FP64, schedule=capi.SCHED_PERSISTENT, step_control=capi.STEP_LEARNED,
                start_mode=capi.START_GAUSSIAN)


def traffic_probe(n, lanes=8):
    """One chosen launch at full size (lanes: path width, 8 or 16 lanes per path), then a 1 GiB
    device copy (counter calibration)."""
    import torch

    from finite_root_generation_amd import capi, tasks
    log("[probe] training the step model")
    step_model, info = train_step_model(capi, tasks)
    log(f"[probe] trained in {info['seconds']} s")
    fpm, gamma = tasks.seeded_task(1)
    dev = torch.device("cuda:0")
    fpt = torch.empty((n, 21), dtype=torch.complex128, device=dev)
    err = torch.empty((n,), dtype=torch.int32, device=dev)
    st = torch.empty((n, 8), dtype=torch.int32, device=dev)
    with capi.Tracker(0, lanes_per_path=lanes, **chosen_opts(capi)) as t:
        t.set_step_model(step_model)
        log(f"[probe] chosen launch, {lanes} lanes per path")
        t.track_device(fpm, gamma, 2, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        torch.cuda.synchronize()
        log(f"[probe] kernel {t.last_kernel_ms():.1f} ms")
    a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    log("[probe] 1 GiB copy")
    b.copy_(a)
    torch.cuda.synchronize()


def shape_of(capi, o):
    names = {capi.SCHED_STATIC: "static", capi.SCHED_PERSISTENT: "persistent"}
    return {"schedule": names[o.schedule], "block_size": o.block_size or 64, "groups_per_cu": o.groups_per_cu,
            "lanes_per_path": o.lanes_per_path or 8, "waves_per_simd": o.waves_per_simd or 1}


def tune_table(capi, rows):
    names = {capi.SCHED_STATIC: "static", capi.SCHED_PERSISTENT: "persistent"}
    return [{"schedule": names[r["schedule"]], "block_size": r["block_size"], "groups_per_cu": r["groups_per_cu"],
             "lanes_per_path": r["lanes_per_path"], "waves_per_simd": r["waves_per_simd"],
             "paths_per_s": round(r["paths_per_s"], 1)} for r in rows]


def batched_run(capi, step_model, fpms, gams, seeds, pids, counts, precision, shape_opts=None):
    """One graph-replayed persistent launch of all instances (warm-up, then timed).  The launch
    shape is autotuned unless shape_opts (an frg_opts with a tuned shape) is given."""
    name = "mixed" if precision == capi.PREC_MIXED else "fp64"
    with capi.Tracker(0, precision=precision, use_graph=1, step_control=capi.STEP_LEARNED,
                      start_mode=capi.START_GAUSSIAN) as t:
        t.set_step_model(step_model)
        rows = None
        if shape_opts is None:
            log(f"[bench] batched {name}: autotune on {TUNE_SAMPLE} paths")
            rows, chosen = t.autotune(fpms[0], gams[0], seed=9, n_sample=TUNE_SAMPLE, first_path_id=1 << 40)
        else:
            t.set_opts(schedule=shape_opts.schedule, block_size=shape_opts.block_size,
                       groups_per_cu=shape_opts.groups_per_cu, lanes_per_path=shape_opts.lanes_per_path,
                       waves_per_simd=shape_opts.waves_per_simd)
        log(f"[bench] batched {name}: shape {shape_of(capi, t.opts)}; warm-up")
        t.track_batch(fpms, gams, seeds, pids, counts)
        log(f"[bench] batched {name}: timed")
        fpt, err, st, summ = t.track_batch(fpms, gams, seeds, pids, counts)
        shape = shape_of(capi, t.opts)
    return fpt, err, summ, shape, rows


def batched(capi, tasks, step_model, with_plain=True, fp64_shape=None):
    """64 instances in one graph-replayed persistent launch with learned step control (see the
    module doc), in mixed precision (autotuned) and in FP64 (the headline's tuned shape); with
    with_plain, also the plain way: each instance its own static FP64 launch, classical control."""
    rng = np.random.default_rng(4)
    counts = np.exp(rng.uniform(np.log(1024), np.log(262144), 64)).astype(np.int64)
    fpms, gams = zip(*[tasks.seeded_task(200 + k) for k in range(64)])
    fpms, gams = np.array(fpms), np.array(gams)
    seeds = 300 + np.arange(64)
    pids = np.zeros(64, np.int64)
    fpt, err, summ, shape, rows = batched_run(capi, step_model, fpms, gams, seeds, pids, counts, capi.PREC_MIXED)
    f64 = batched_run(capi, step_model, fpms, gams, seeds, pids, counts, capi.PREC_FP64, shape_opts=fp64_shape)
    n = int(counts.sum())
    plain = None
    if with_plain:
        log("[bench] batched: plain, one static FP64 launch per instance")
        plain_ms = 0.0
        pf, pe = [], []
        with capi.Tracker(0, start_mode=capi.START_GAUSSIAN) as t:
            for k in range(64):
                f, e, _, s = t.track(fpms[k], gams[k], int(seeds[k]), int(counts[k]), want_stats=True)
                plain_ms += s.kernel_ms + s.refine_ms
                pf.append(f)
                pe.append(e)
                if k % 8 == 7:
                    log(f"[bench] batched: plain, {k + 1}/64 instances")
        pf, pe = np.concatenate(pf), np.concatenate(pe)
        log("[bench] batched: distinct roots per instance and recall against the plain launches")
        ok_c, ok_p = tasks.finite_mask(fpt, err), tasks.finite_mask(pf, pe)
        both = ok_c & ok_p
        same_path = both & (np.max(np.abs(fpt - pf), axis=1) <= ROOT_TOL)
        offs = np.concatenate([[0], np.cumsum(counts)])
        sets_p = instance_sets(capi, tasks, pf, pe, offs)
        sets_m = instance_sets(capi, tasks, fpt, err, offs)
        sets_f = instance_sets(capi, tasks, f64[0], f64[1], offs)
        plain = {"value": round(n / (plain_ms * 1e-3), 1), "kernel_ms": round(plain_ms, 1),
                 "how": "64 separate launches, static schedule, FP64, classical step control",
                 "finite_roots": int(ok_p.sum()), "same_root_on_same_path": int(same_path.sum()),
                 "distinct_roots": int(sum(len(x) for x in sets_p))}
        set_stats = {"mixed": {"distinct_roots": int(sum(len(x) for x in sets_m)),
                               "recall_of_plain_sets": round(sets_recall(sets_p, sets_m), 6),
                               "recall_by_plain_sets": round(sets_recall(sets_m, sets_p), 6)},
                     "fp64": {"distinct_roots": int(sum(len(x) for x in sets_f)),
                              "recall_of_plain_sets": round(sets_recall(sets_p, sets_f), 6),
                              "recall_by_plain_sets": round(sets_recall(sets_f, sets_p), 6)},
                     "how": "per instance, the reference's dedupe (1e-6), then roots matched at 1e-8 absolute"}
    ms = summ.kernel_ms + summ.refine_ms
    f_fpt, f_err, f_summ, f_shape, _ = f64
    f_ms = f_summ.kernel_ms + f_summ.refine_ms
    f_ok = tasks.finite_mask(f_fpt, f_err)
    fp64 = {"value": round(n / (f_ms * 1e-3), 1), "launch_ms": round(f_ms, 1), "precision": "fp64",
            "step_control": "learned", "graph": True, "shape": f_shape,
            "finite_roots": int(f_ok.sum()), "failures": np.bincount(f_err, minlength=6).tolist()}
    if plain:
        fp64.update(set_stats["fp64"])
        fp64["speedup_vs_plain"] = round(fp64["value"] / plain["value"], 3)
        fp64["same_root_on_same_path_as_plain"] = int((f_ok & tasks.finite_mask(pf, pe) & (
            np.max(np.abs(f_fpt - pf), axis=1) <= ROOT_TOL)).sum())
    return {"instances": 64, "paths": n, "path_counts": {"min": int(counts.min()), "max": int(counts.max())},
            "value": round(n / (ms * 1e-3), 1), "unit": "paths/s", "launch_ms": round(ms, 1),
            "pcie_inclusive_paths_per_s": round(n / (summ.total_ms * 1e-3), 1),
            "precision": "fp32+fp64 mixed", "step_control": "learned", "graph": True,
            "autotune": {"chosen": shape, "sample_paths": TUNE_SAMPLE, "candidates": tune_table(capi, rows)},
            "finite_roots": int(tasks.finite_mask(fpt, err).sum()),
            "failures": np.bincount(err, minlength=6).tolist(),
            "plain": plain, "speedup_vs_plain": round(n / (ms * 1e-3) / plain["value"], 3) if plain else None,
            "distinct_roots": set_stats["mixed"] if plain else None, "fp64_learned": fp64}


def cpu_baseline(fpm, gamma, seed, n_sample, start_mode):
    """The oracle (the CPU restatement of the reference's per-path routines) on the first n_sample
    paths of the same workload, on every core of the process affinity; returns the timing and the
    oracle's outputs for the per-path comparison."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    oo = O.default_opts(start_mode=start_mode)
    t0 = time.perf_counter()
    fpt, err, _, _ = O.homotopy(oo, fpm, gamma, seed, 0, n_sample, nthreads=cores, fast=True)
    dt = time.perf_counter() - t0
    n_one = max(16, n_sample // (8 * cores))
    t0 = time.perf_counter()
    O.homotopy(oo, fpm, gamma, seed, 0, n_one, nthreads=1, fast=True)
    dt1 = time.perf_counter() - t0
    res = {"value": round(n_sample / dt, 2), "unit": "paths/s", "cores": cores, "kind": "port",
           "nproc": os.cpu_count(),
           "sample": f"first {n_sample} paths of the same workload (same task, seeds and RNG ids), "
                     f"CPU oracle -O3, OpenMP {cores} threads (every core of the process affinity), {dt:.1f} s",
           "single_core": {"value": round(n_one / dt1, 2), "unit": "paths/s",
                           "sample": f"first {n_one} paths, one thread, {dt1:.1f} s"}}
    return res, fpt, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=0, help="timed plain launches (0: min(steps, 5))")
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="paths for the CPU baseline (0: 1000 per core, about 10 s of work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ablations", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-batched-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="chosen launch shape without the autotuner")
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--probe-lanes", type=int, default=8, choices=(8, 16), help="path width of the traffic probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; more GPUs would each track a contiguous "
                                   "range of path ids, which is not launched here", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.traffic_probe:
        traffic_probe(args.paths, args.probe_lanes)
        return

    import torch

    from finite_root_generation_amd import capi, tasks

    n = args.paths
    fpm, gamma = tasks.seeded_task(1)
    seed = 2
    consts = flop_constants()
    log("[bench] training the step model")
    step_model, train_info = train_step_model(capi, tasks)
    plain_steps = args.plain_steps or min(args.steps, 5)
    plain = run_variant(capi, torch, "plain", dict(precision=capi.PREC_FP64, schedule=capi.SCHED_STATIC,
                                                   start_mode=capi.START_GAUSSIAN),
                        fpm, gamma, n, plain_steps, min(args.warmup, 1), seed)
    chosen = run_variant(capi, torch, "chosen", chosen_opts(capi), fpm, gamma, n, args.steps, args.warmup, seed,
                         model=step_model, tune=None if args.no_tune else (9, min(TUNE_SAMPLE, n)))
    ablations = {}
    if not args.no_ablations:
        co = chosen["opts"]
        variants = [
            ("persistent_classical", dict(precision=capi.PREC_FP64, schedule=capi.SCHED_PERSISTENT), None),
            ("mixed_persistent_classical", dict(precision=capi.PREC_MIXED, schedule=capi.SCHED_PERSISTENT), None),
            ("mixed_persistent_learned", dict(precision=capi.PREC_MIXED, schedule=capi.SCHED_PERSISTENT,
                                              step_control=capi.STEP_LEARNED), step_model)]
        if co.lanes_per_path == 16:
            # the chosen controls with 8 lanes per path: what the 16-lane kernel adds
            variants.append(("persistent_learned_8_lanes", dict(precision=capi.PREC_FP64,
                                                                schedule=capi.SCHED_PERSISTENT,
                                                                step_control=capi.STEP_LEARNED), step_model))
        for name, kw, mdl in variants:
            o = dict(kw, start_mode=capi.START_GAUSSIAN)
            # the chosen lanes per path where the precision allows it (the 16-lane solve is FP64)
            if co.lanes_per_path == 16 and kw["precision"] == capi.PREC_FP64 and "8_lanes" not in name:
                o["lanes_per_path"] = 16
                o["waves_per_simd"] = co.waves_per_simd
            r = run_variant(capi, torch, name, o, fpm, gamma, n, 1, 0, seed, model=mdl)
            ok_r, ok_p = tasks.finite_mask(r["fpt"], r["err"]), tasks.finite_mask(plain["fpt"], plain["err"])
            ablations[name] = {"value": round(n / (r["ms_per_step"] * 1e-3), 1),
                               "rk_attempts_per_path": round(float(r["stats"]["rk_attempts"].mean()), 3),
                               "finite_roots": int(ok_r.sum()),
                               "same_root_on_same_path_as_plain": int((ok_r & ok_p & (np.max(np.abs(
                                   r["fpt"] - plain["fpt"]), axis=1) <= ROOT_TOL)).sum()),
                               "failures": np.bincount(r["err"], minlength=6).tolist()}
    log("[bench] distinct roots (reference dedupe rule, tolerance 1e-6)")
    set_c, dt_c = distinct_roots(capi, chosen)
    set_p, dt_p = distinct_roots(capi, plain)
    log("[bench] set-level recall")
    recall_p_in_c = set_recall(set_p, set_c)
    recall_c_in_p = set_recall(set_c, set_p)
    value = n / (chosen["ms_per_step"] * 1e-3)
    plain_value = n / (plain["ms_per_step"] * 1e-3)
    lengths = plain["stats"]["steps"]
    hist, edges = np.histogram(lengths, bins=[0, 25, 50, 75, 100, 150, 200, 300, 500, 1000, 2001])
    shape = shape_of(capi, chosen["opts"])
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(chosen["ms_per_step"], 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic: seeded generic task (seed 1), Philox start systems (seed 2, standard complex Gaussian "
                "free draws)",
        "config": {"workload": "stephenson2_frg_full_size", "paths": n, "task_seed": 1, "start_seed": 2,
                   "start_draws": "standard complex Gaussian",
                   "chosen": "FP64, persistent work queue, learned step control, autotuned launch shape",
                   "chosen_launch_shape": shape,
                   "plain": "FP64, static grid, classical step control, 8 lanes per path, 64-thread blocks"},
        "roofline": roofline(chosen, consts, traffic(chosen["kernel_ms"])),
        "step_breakdown_ms": {"tracking_kernel": round(chosen["kernel_ms"], 2),
                              "refinement_kernel": round(chosen["refine_ms"], 2)},
        "host_window": {"d2h_ms": round(chosen["d2h_ms"], 1),
                        "paths_per_s": round(n / ((chosen["ms_per_step"] + chosen["d2h_ms"]) * 1e-3), 1),
                        "how": "launch -> roots and codes on the host (the reference's window, frg.cu:356-426)"},
        "plain": {"value": round(plain_value, 1), "ms_per_step": round(plain["ms_per_step"], 2),
                  "steps": plain_steps, "roofline": roofline(plain, consts),
                  "host_window_paths_per_s": round(n / ((plain["ms_per_step"] + plain["d2h_ms"]) * 1e-3), 1)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "rk_attempts_per_path": {"chosen": round(float(chosen["stats"]["rk_attempts"].mean()), 3),
                                 "plain": round(float(plain["stats"]["rk_attempts"].mean()), 3)},
        "solves_per_path": {k: round(float((r["stats"]["tangent_solves"] + r["stats"]["newton_its"]).mean()), 2)
                            for k, r in (("chosen", chosen), ("plain", plain))},
        "step_model": train_info,
        "ablations": ablations,
        "finite_roots": {"chosen": int(tasks.finite_mask(chosen["fpt"], chosen["err"]).sum()),
                         "plain": int(tasks.finite_mask(plain["fpt"], plain["err"]).sum())},
        "distinct_roots": {"chosen": int(len(set_c)), "plain": int(len(set_p)),
                           "dedupe_tolerance": DEDUPE_TOL,
                           "chosen_per_s": round(len(set_c) / (chosen["ms_per_step"] * 1e-3), 1),
                           "plain_per_s": round(len(set_p) / (plain["ms_per_step"] * 1e-3), 1),
                           "recall_of_plain_set_by_chosen": round(recall_p_in_c, 6),
                           "recall_of_chosen_set_by_plain": round(recall_c_in_p, 6),
                           "match_tolerance": ROOT_TOL, "host_dedupe_s": round(dt_c + dt_p, 2)},
        "failures": {"chosen": np.bincount(chosen["err"], minlength=6).tolist(),
                     "plain": np.bincount(plain["err"], minlength=6).tolist()},
        "path_length_hist": {"edges": edges.tolist(), "counts": hist.tolist()},
        "autotune": None if chosen["tuned"] is None else {"sample_paths": min(TUNE_SAMPLE, n),
                                                          "candidates": tune_table(capi, chosen["tuned"]["rows"])},
        "batched": None,
        "cpu_baseline": None,
        "oracle_comparison": None,
    }
    if not args.no_batched:
        line["batched"] = batched(capi, tasks, step_model, with_plain=not args.no_batched_plain,
                                  fp64_shape=chosen["opts"])
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        ns = min(n, args.cpu_sample or max(64, 1000 * cores))
        log(f"[bench] cpu baseline: {ns} paths on {cores} cores")
        cb, o_fpt, o_err = cpu_baseline(fpm, gamma, seed, ns, capi.START_GAUSSIAN)
        line["cpu_baseline"] = cb
        log("[bench] per-path comparison with the oracle sample")
        line["oracle_comparison"] = {
            "plain": per_path_vs_oracle(tasks, plain["fpt"], plain["err"], o_fpt, o_err),
            "chosen": per_path_vs_oracle(tasks, chosen["fpt"], chosen["err"], o_fpt, o_err),
            "note": "same path ids 0..n-1; plain follows the reference's step sequence, so its paths agree one by "
                    "one except where rounding decides the end point; the chosen tracker's learned step control "
                    "changes the step sequence, so its comparison is by codes and finite roots",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
