"""Benchmark: tracked paths/s on one MI355X, chosen tracker vs plain FP64, at equal recall.

Workload (one "step" = one pass over one batch): the Stephenson II finite-root-generation round
at full size, 1,048,576 paths; target task from seed 1 (finite_root_generation_amd.tasks
.seeded_task), start systems drawn by the tracker's Philox stream from seed 2.  Inputs are the
30 target parameters and the seed; outputs (roots, codes, per-path counters) stay in HBM, so
the timed region is the tracking launch alone.  The same batch is run by
  * plain:  FP64, classical step control, one path per 8-lane group, no refill (the
            reference's launch shape);
  * chosen: FP64, persistent grid with a device work queue, learned step-size control (model
            trained before the timed region on step logs of other tasks: seeds 10, 20, 30);
  * ablations (one launch each): persistent + classical, mixed FP32/FP64 + persistent + classical,
    and mixed + persistent + learned.
value = chosen paths/s.  One JSON line on stdout.

The line also carries "batched", the many-instance workload: 64 task instances (seeded tasks
200..263, start seeds 300..363) with path counts log-uniform in [1,024, 262,144] (numpy seed 4),
FP32/FP64 mixed precision and learned step control, launch shape picked by the autotuner on a
65,536-path sample, all instances in one persistent launch replayed from a hipGraph; next to it the
plain way, each instance its own static FP64 launch with classical control, as the reference runs
one target at a time.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--paths N] [--cpu-sample N]
  python bench.py --traffic-probe     (for rocprofv3 --pmc runs: one chosen launch, then a plain
                                       copy of 1 GiB whose counter readings calibrate the counters)

roofline.traffic comes from profiles/track_kernel_traffic.json when that file is present: HBM
bytes per chosen launch measured by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of the
traffic probe, each scaled by what the counter reads for the 1 GiB copy.
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
P64_TFLOPS = 78.6    # MI355X FP64 vector/matrix peak
P32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_variant(capi, torch, name, opts_kw, fpm, gamma, n, steps, warmup, seed, model=None):
    t = capi.Tracker(0, **opts_kw)
    if model is not None:
        t.set_step_model(model)
    dev = torch.device("cuda:0")
    fpt = torch.empty((n, 21), dtype=torch.complex128, device=dev)
    err = torch.empty((n,), dtype=torch.int32, device=dev)
    st = torch.empty((n, 8), dtype=torch.int32, device=dev)
    for w in range(warmup):
        log(f"[bench] {name}: warmup {w + 1}/{warmup}")
        t.track_device(fpm, gamma, seed, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        torch.cuda.synchronize()
    kern = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        t.track_device(fpm, gamma, seed, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        kern.append(t.last_kernel_ms())
        log(f"[bench] {name}: step {k + 1}/{steps} kernel {kern[-1]:.1f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        "fpt": fpt.cpu().numpy(), "err": err.cpu().numpy(),
        "stats": st.cpu().numpy().view(capi.PATH_STATS_DTYPE).reshape(n),
        "ms_per_step": 1e3 * wall / steps, "kernel_ms": float(np.mean(kern)),
    }
    t.close()
    return out


def flops_split(capi, stats, model):
    """(FP64 flops, FP32 flops) of the tracked paths: FP32 solves counted separately."""
    solves = stats["tangent_solves"].astype(np.float64) + stats["newton_its"].astype(np.float64)
    f32_share = np.where(solves > 0, stats["fp32_solves"] / np.maximum(solves, 1), 0.0)
    per_solve = (stats["tangent_solves"] * model["tangent_solve"] + stats["newton_its"] * model["newton_update"])
    f32 = float((per_solve * f32_share).sum())
    f64 = float((per_solve * (1 - f32_share)).sum() + stats["rk_attempts"].astype(np.float64).sum()
                * model["rk_combination"])
    return f64, f32


def roofline(capi, res, model, hbm=None):
    f64, f32 = flops_split(capi, res["stats"], model)
    sec = res["kernel_ms"] * 1e-3
    tmin = f64 / (P64_TFLOPS * 1e12) + f32 / (P32_TFLOPS * 1e12)
    achieved = (f64 + f32) / sec / 1e12
    peak = (f64 + f32) / tmin / 1e12 if tmin > 0 else P64_TFLOPS
    # compute-bound ("mfma" in the line's vocabulary, against HBM); the arithmetic runs on the FP64
    # vector ALU, no MFMA instruction is used, and MI355X's FP64 peak (78.6 TFLOP/s) is the same
    # for the vector and the matrix units
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": hbm["bytes"] if hbm else None,
            "bound_detail": "FP64 vector ALU (no MFMA instructions); peak = MI355X FP64 78.6 TFLOP/s",
            "traffic_detail": hbm, "flops_per_launch": {"fp64": f64, "fp32": f32}}


def compare(tasks, a, b, tol=1e-8):
    """Recall of b's finite roots by a: a root of b counts as found when a has a finite root
    within tol (max over components of |dx| / max(1, |x|)) -- on the same path, or else anywhere
    (candidates from a window of +-tol(1 + |Re z0|) on a's sorted Re z0, then every component)."""
    fa = tasks.finite_mask(a["fpt"], a["err"])
    fb = tasks.finite_mask(b["fpt"], b["err"])
    both = fa & fb
    d = np.full(len(fa), np.inf)
    d[both] = np.max(np.abs(a["fpt"][both] - b["fpt"][both]) / np.maximum(1.0, np.abs(b["fpt"][both])), axis=1)
    match = both & (d <= tol)
    found = int(match.sum())
    rest = np.nonzero(fb & ~match)[0]
    if len(rest) and fa.any():
        ra = a["fpt"][fa]
        order = np.argsort(ra[:, 0].real, kind="stable")
        keys = ra[order, 0].real
        for i in rest:
            x = b["fpt"][i]
            w = tol * max(1.0, abs(x[0]))
            lo, hi = np.searchsorted(keys, [x[0].real - w, x[0].real + w], side="left")[0], \
                np.searchsorted(keys, x[0].real + w, side="right")
            if hi > lo:
                c = ra[order[lo:hi]]
                if np.any(np.max(np.abs(c - x) / np.maximum(1.0, np.abs(x)), axis=1) <= tol):
                    found += 1
    return {"finite_a": int(fa.sum()), "finite_b": int(fb.sum()), "per_path_match": int(match.sum()),
            "recall": float(found / max(1, fb.sum()))}


def train_step_model(capi, tasks):
    """Step logs of three other tasks (seeds 10, 20, 30), one model trained on the GPU."""
    t0 = time.perf_counter()
    recs = []
    with capi.Tracker(0) as t:
        for k, s in enumerate((10, 20, 30)):
            fpm, g = tasks.seeded_task(s)
            recs.append(t.collect_step_log(fpm, g, seed=100 + k, n=32768, cap=4_000_000))
        rec = np.concatenate(recs)
        rec = rec[: 10_000_000]
        model, loss = t.train_step_model(rec, steps=4000, batch=4096, lr=2e-3, seed=1)
    return model, {"records": int(rec.shape[0]), "train_steps": 4000, "final_loss": float(loss[-50:].mean()),
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


def traffic_probe(n):
    """One chosen launch at full size, then a 1 GiB device copy (counter calibration)."""
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
    with capi.Tracker(0, precision=capi.PREC_FP64, schedule=capi.SCHED_PERSISTENT,
                      step_control=capi.STEP_LEARNED) as t:
        t.set_step_model(step_model)
        log("[probe] chosen launch")
        t.track_device(fpm, gamma, 2, n, fpt.data_ptr(), err.data_ptr(), st.data_ptr())
        torch.cuda.synchronize()
        log(f"[probe] kernel {t.last_kernel_ms():.1f} ms")
    a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    log("[probe] 1 GiB copy")
    b.copy_(a)
    torch.cuda.synchronize()


def batched(capi, tasks, step_model, with_plain=True):
    """64 instances in one graph-replayed persistent launch, autotuned (see the module doc); with
    with_plain, also the plain way: each instance its own static FP64 launch, classical control."""
    rng = np.random.default_rng(4)
    counts = np.exp(rng.uniform(np.log(1024), np.log(262144), 64)).astype(np.int64)
    fpms, gams = zip(*[tasks.seeded_task(200 + k) for k in range(64)])
    fpms, gams = np.array(fpms), np.array(gams)
    seeds = 300 + np.arange(64)
    pids = np.zeros(64, np.int64)
    with capi.Tracker(0, precision=capi.PREC_MIXED, use_graph=1, step_control=capi.STEP_LEARNED) as t:
        t.set_step_model(step_model)
        log("[bench] batched: autotune")
        rows, chosen = t.autotune(fpms[0], gams[0], seed=9, n_sample=65536, first_path_id=1 << 40)
        log(f"[bench] batched: chosen schedule/block/groups-per-CU {chosen}; warm-up")
        t.track_batch(fpms, gams, seeds, pids, counts)
        log("[bench] batched: timed")
        fpt, err, st, summ = t.track_batch(fpms, gams, seeds, pids, counts)
    n = int(counts.sum())
    plain = None
    if with_plain:
        log("[bench] batched: plain, one static FP64 launch per instance")
        plain_ms = 0.0
        pf, pe = [], []
        with capi.Tracker(0) as t:
            for k in range(64):
                f, e, _, _ = t.track(fpms[k], gams[k], int(seeds[k]), int(counts[k]), want_stats=False)
                plain_ms += t.last_kernel_ms()
                pf.append(f)
                pe.append(e)
                if k % 8 == 7:
                    log(f"[bench] batched: plain, {k + 1}/64 instances")
        log("[bench] batched: recall against the plain launches")
        c = compare(tasks, {"fpt": fpt, "err": err}, {"fpt": np.concatenate(pf), "err": np.concatenate(pe)})
        plain = {"value": round(n / (plain_ms * 1e-3), 1), "kernel_ms": round(plain_ms, 1),
                 "how": "64 separate launches, static schedule, FP64, classical step control",
                 "finite_roots": c["finite_b"], "recall_of_plain_roots": round(c["recall"], 6)}
    names = {capi.SCHED_STATIC: "static", capi.SCHED_PERSISTENT: "persistent"}
    return {"instances": 64, "paths": n, "path_counts": {"min": int(counts.min()), "max": int(counts.max())},
            "value": round(n / (summ.kernel_ms * 1e-3), 1), "unit": "paths/s", "kernel_ms": round(summ.kernel_ms, 1),
            "pcie_inclusive_paths_per_s": round(n / (summ.total_ms * 1e-3), 1),
            "precision": "fp32+fp64 mixed", "step_control": "learned", "graph": True,
            "autotune": {"chosen": {"schedule": names[chosen[0]], "block_size": chosen[1], "groups_per_cu": chosen[2]},
                         "sample_paths": 65536,
                         "candidates": [{"schedule": names[r["schedule"]], "block_size": r["block_size"],
                                         "groups_per_cu": r["groups_per_cu"], "paths_per_s": round(r["paths_per_s"], 1)}
                                        for r in rows]},
            "finite_roots": int(tasks.finite_mask(fpt, err).sum()),
            "failures": np.bincount(err, minlength=6).tolist(),
            "plain": plain, "speedup_vs_plain": round(n / (summ.kernel_ms * 1e-3) / plain["value"], 3) if plain else None}


def cpu_baseline(fpm, gamma, seed, n_sample):
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    cores = max(1, min(cores, 16))
    oo = O.default_opts()
    t0 = time.perf_counter()
    O.homotopy(oo, fpm, gamma, seed, 0, n_sample, nthreads=cores, fast=True)
    dt = time.perf_counter() - t0
    n_one = max(16, n_sample // (8 * cores))
    t0 = time.perf_counter()
    O.homotopy(oo, fpm, gamma, seed, 0, n_one, nthreads=1, fast=True)
    dt1 = time.perf_counter() - t0
    return {"value": round(n_sample / dt, 2), "unit": "paths/s", "cores": cores, "kind": "port",
            "sample": f"first {n_sample} paths of the same workload (same task, seeds and RNG ids), "
                      f"CPU oracle -O3, OpenMP {cores} threads, {dt:.1f} s",
            "single_core": {"value": round(n_one / dt1, 2), "unit": "paths/s",
                            "sample": f"first {n_one} paths, one thread, {dt1:.1f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="paths for the CPU baseline (0: 1000 per thread, about 10 s of work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ablations", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-batched-plain", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; more GPUs would each track a contiguous "
                                   "range of path ids, which is not launched here", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.traffic_probe:
        traffic_probe(args.paths)
        return

    import torch

    from finite_root_generation_amd import capi, tasks

    n = args.paths
    fpm, gamma = tasks.seeded_task(1)
    seed = 2
    model = capi.flop_model()
    log("[bench] training the step model")
    step_model, train_info = train_step_model(capi, tasks)
    plain = run_variant(capi, torch, "plain", dict(precision=capi.PREC_FP64, schedule=capi.SCHED_STATIC),
                        fpm, gamma, n, args.steps, args.warmup, seed)
    chosen = run_variant(capi, torch, "chosen", dict(precision=capi.PREC_FP64, schedule=capi.SCHED_PERSISTENT,
                                                     step_control=capi.STEP_LEARNED),
                         fpm, gamma, n, args.steps, args.warmup, seed, model=step_model)
    ablations = {}
    if not args.no_ablations:
        for name, kw, mdl in (
                ("persistent_classical", dict(precision=capi.PREC_FP64, schedule=capi.SCHED_PERSISTENT), None),
                ("mixed_persistent_classical", dict(precision=capi.PREC_MIXED, schedule=capi.SCHED_PERSISTENT), None),
                ("mixed_persistent_learned", dict(precision=capi.PREC_MIXED, schedule=capi.SCHED_PERSISTENT,
                                                  step_control=capi.STEP_LEARNED), step_model)):
            r = run_variant(capi, torch, name, kw, fpm, gamma, n, 1, 0, seed, model=mdl)
            c = compare(tasks, r, plain)
            ablations[name] = {"value": round(n / (r["ms_per_step"] * 1e-3), 1),
                               "rk_attempts_per_path": round(float(r["stats"]["rk_attempts"].mean()), 3),
                               "recall_vs_plain": round(c["recall"], 6), "finite_roots": c["finite_a"],
                               "failures": np.bincount(r["err"], minlength=6).tolist()}
    batch = None if args.no_batched else batched(capi, tasks, step_model, with_plain=not args.no_batched_plain)
    value = n / (chosen["ms_per_step"] * 1e-3)
    plain_value = n / (plain["ms_per_step"] * 1e-3)
    cmp = compare(tasks, chosen, plain)
    lengths = plain["stats"]["steps"]
    hist, edges = np.histogram(lengths, bins=[0, 25, 50, 75, 100, 150, 200, 300, 500, 1000, 2001])
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(chosen["ms_per_step"], 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic: seeded generic task (seed 1), Philox start systems (seed 2)",
        "config": {"workload": "stephenson2_frg_full_size", "paths": n, "task_seed": 1, "start_seed": 2,
                   "chosen": "FP64, persistent work queue, learned step control",
                   "plain": "FP64, static, classical step control"},
        "roofline": roofline(capi, chosen, model, traffic(chosen["kernel_ms"])),
        "plain": {"value": round(plain_value, 1), "ms_per_step": round(plain["ms_per_step"], 2),
                  "roofline": roofline(capi, plain, model)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "rk_attempts_per_path": {"chosen": round(float(chosen["stats"]["rk_attempts"].mean()), 3),
                                 "plain": round(float(plain["stats"]["rk_attempts"].mean()), 3)},
        "step_model": train_info,
        "ablations": ablations,
        "finite_roots": {"chosen": cmp["finite_a"], "plain": cmp["finite_b"]},
        "failures": {"chosen": np.bincount(chosen["err"], minlength=6).tolist(),
                     "plain": np.bincount(plain["err"], minlength=6).tolist()},
        "recall_vs_plain": round(cmp["recall"], 6),
        "path_length_hist": {"edges": edges.tolist(), "counts": hist.tolist()},
        "batched": batch,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        ns = args.cpu_sample or max(64, 1000 * min(16, cores))
        log(f"[bench] cpu baseline: {ns} paths")
        line["cpu_baseline"] = cpu_baseline(fpm, gamma, seed, ns)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
