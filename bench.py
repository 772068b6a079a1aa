"""Benchmark: eigenpairs/s of the batched MRRR solve against the plain one-matrix-per-
solve path on the mixed-size workload, on one MI355X.

Workload (the batched mixed workload): 4096 FP64 tridiagonals, n log-uniform in
[32, 2048], 50% uniform(-1,1), 25% perturbed 1-2-1 Toeplitz, 25% glued Wilkinson
(glue 1e-10), jobz = V, range = A, synthetic data from fixed seeds.  Inputs and
outputs are resident in HBM before the timed region starts.

Prints one JSON line.  Besides the contract fields it reports the plain path's
rate and the ratio, a short served stream (p50/p99 latency), resid/orth of a
sample, per-kernel times from HIP events recorded on the library's streams during
the timed steps, the roofline of the dominant kernel (algorithmic FLOPs counted on
the device in one extra step), and the CPU baseline (the compiled reference on the host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

# FP64 operations per counted element, per kernel family (algorithmic work:
# Sturm step sub/div/sub = 3, dlaneg add/div/mul/sub = 4, dlar1v 16 per element,
# dlarrf sweep 5 per element).
# The small-matrix route (k_small, n < 500, dstemr): dqds step 5 (add, divide, two
# multiplies, subtract), dlaneg 4, Sturm step on T 3, dlar1v row 16, dlarrf element 5.
FLOPS_PER_STEP = {"coarse": 3, "refine": 4, "endref": 4, "r1v": 16, "r1v_bisect": 4,
                  "rrf": 5, "rrj": 3, "refine_cl": 4,
                  "sm_dqds": 5, "sm_neg": 4, "sm_sturm": 3, "sm_r1v": 16, "sm_rrf": 5,
                  "tail_r1v": 16, "tail_bisect": 4, "tailb_r1v": 16, "tailb_bisect": 4}
KERNEL_STEPS = {"k_coarse": ["coarse"], "k_refine_root": ["refine"], "k_endref": ["endref"],
                "k_s_iter": ["r1v"], "k_s_bisect": ["r1v_bisect"], "k_dlarrf": ["rrf"], "k_rrj": ["rrj"],
                "k_refine_cl": ["refine_cl"],
                "k_small": ["sm_dqds", "sm_neg", "sm_sturm", "sm_r1v", "sm_rrf"],
                "k_s_tail<false>": ["tail_r1v", "tail_bisect"], "k_s_tail<true>": ["tailb_r1v", "tailb_bisect"]}
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# The divide-bound ceiling: the IEEE FP64 division sequence the compiler emits is ten
# VALU instructions (v_div_scale x2, v_rcp_f64, four FMAs, v_mul, v_div_fmas,
# v_div_fixup); at 32 FP64 flops per cycle per SIMD (78.6 TF / (256 CUs x 4 SIMDs x
# 2.4 GHz)) a wave64 FP64 instruction takes 4 cycles and the quarter-rate v_rcp_f64 16,
# so one division issues in 52 cycles per wave64 on one SIMD.  An upper bound for any
# loop with one division per step (it leaves out the step's other instructions):
# 256 CUs x 4 SIMDs x 2.4 GHz x 64 lanes / 52 cycles division steps per second.
DIVIDE_STEPS_PER_S = 256 * 4 * 2.4e9 * 64 / 52.0
# divisions the reference's algorithm needs per counted step of each family (dlar1v: the
# stationary and the progressive transform; the vector solve multiplies)
DIVIDES_PER_STEP = {"coarse": 1, "refine": 1, "endref": 1, "r1v": 2, "r1v_bisect": 1,
                    "rrf": 1, "rrj": 1, "refine_cl": 1,
                    "sm_dqds": 1, "sm_neg": 1, "sm_sturm": 1, "sm_r1v": 2, "sm_rrf": 1,
                    "tail_r1v": 2, "tail_bisect": 1, "tailb_r1v": 2, "tailb_bisect": 1}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(torch, dev, count):
    from mr3smp_amd import workloads as WL
    mats = WL.batched_mixed(count)
    probs = []
    for D, E, kind in mats:
        n = len(D)
        probs.append(dict(D=torch.tensor(D, device=dev), E=torch.tensor(E, device=dev),
                          jobz="V", range="A",
                          W=torch.empty(n, dtype=torch.float64, device=dev),
                          Z=torch.empty((n, n), dtype=torch.float64, device=dev),
                          Zsupp=torch.empty(2 * n, dtype=torch.int32, device=dev)))
    return mats, probs


def accuracy_sample(mats, keep, arr, stride):
    from mrrr_ctypes import resid_orth
    worst_r, worst_o = 0.0, 0.0
    for i in range(0, len(mats), stride):
        D, E, _ = mats[i]
        m = arr[i].m
        W = keep[i]["W"][:m].cpu().numpy()
        Z = keep[i]["Z"][:m].cpu().numpy().T
        r, o = resid_orth(D, E, W, Z)
        worst_r, worst_o = max(worst_r, r), max(worst_o, o)
    return worst_r, worst_o


def cpu_baseline(mats, stride, threads):
    """The compiled reference (oracle/_ref) when present, else the oracle, on every
    stride-th matrix of the batched workload, run the two ways a CPU user would run it:
    'callers' -- `threads` concurrent callers with PMR_NUM_THREADS=1 (every matrix takes
    the reference's sequential dstemr route; the routine is reentrant); 'threads' -- one
    matrix at a time with PMR_NUM_THREADS=threads (the reference's own routing: dstemr
    below n = 500, the task-parallel route above).  value is the better of the two (the
    oracle, being sequential, has the first only)."""
    from concurrent.futures import ThreadPoolExecutor
    from mrrr_ctypes import ORACLE_SO, REF_SO, call_mrrr, load
    if os.path.exists(REF_SO):
        _, fn = load(REF_SO, "mrrr")
        kind = "reference"
    elif os.path.exists(ORACLE_SO):
        _, fn = load(ORACLE_SO, "oracle_mrrr")
        kind = "port"
    else:
        return None
    sample = mats[::stride]
    pairs = sum(len(D) for D, _, _ in sample)

    def one(t):
        D, E, _ = t
        r = call_mrrr(fn, "V", "A", D, E)
        return r["m"]

    modes = {}
    os.environ["PMR_NUM_THREADS"] = "1"
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        ms = list(ex.map(one, sample))
    dt = time.perf_counter() - t0
    assert sum(ms) == pairs
    modes["callers"] = dict(value=pairs / dt, wall_s=dt,
                            how=f"{threads} concurrent callers, PMR_NUM_THREADS=1")
    if kind == "reference":
        os.environ["PMR_NUM_THREADS"] = str(threads)
        t0 = time.perf_counter()
        ms = [one(t) for t in sample]
        dt = time.perf_counter() - t0
        assert sum(ms) == pairs
        modes["threads"] = dict(value=pairs / dt, wall_s=dt,
                                how=f"one matrix at a time, PMR_NUM_THREADS={threads}")
    os.environ.pop("PMR_NUM_THREADS", None)
    best = max(modes, key=lambda k: modes[k]["value"])
    return dict(value=modes[best]["value"], unit="eigenpairs/s", cores=threads, kind=kind,
                mode=best, modes=modes,
                sample=f"every {stride}th matrix of the batched workload ({len(sample)} matrices, "
                       f"{pairs} eigenpairs, jobz=V range=A); value: the better mode ({best})")


def timed_steps(M, torch, arr, stream, steps, warmup, pairs):
    """Warmup solves, then `steps` timed mr3_solve_batch calls over the same problem
    array, then one solve with the device work counters on.  Returns (seconds of the
    timed steps, their per-kernel report, the counted step's report)."""
    for _ in range(warmup):
        rc = M.run_batch(arr, stream=stream)
        assert rc == 0
    # timed steps: every launch bracketed by HIP events on its own stream (timing only,
    # the kernels run exactly as unprofiled)
    M.profile_reset()
    M.profile_enable(True, counters=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rc = M.run_batch(arr, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    M.profile_enable(False)
    assert rc == 0 and sum(a.m for a in arr) == pairs
    log(f"batched: {dt / steps * 1e3:.1f} ms/step, {pairs * steps / dt:.4g} eigenpairs/s")
    prof = M.profile_report()
    # one more step with the device work counters on: the algorithmic work of a step
    # (the counts do not depend on timing, so every step has the same)
    M.profile_reset()
    M.profile_enable(True)
    rc = M.run_batch(arr, stream=stream)
    torch.cuda.synchronize()
    M.profile_enable(False)
    assert rc == 0
    counted = M.profile_report()
    M.profile_reset()
    return dt, prof, counted


def graph_steps(M, torch, arr, stream, steps, pairs):
    """The same batch captured once as one HIP graph (mr3_graph_create: the small
    route and the pipeline with the tree's depth fixed, in memory the graph owns) and
    replayed `steps` times, each replay waited for like a mr3_solve_batch call.  Returns
    (seconds of the replays, seconds of the capture)."""
    t0 = time.perf_counter()
    g = M.BatchGraph(arr, stream=stream)
    tcap = time.perf_counter() - t0
    assert g.rc == 0, g.rc
    assert g.launch(stream) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rc = g.launch(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert rc == 0 and sum(a.m for a in arr) == pairs
    g.close()
    log(f"graph: {dt / steps * 1e3:.1f} ms/step, {pairs * steps / dt:.4g} eigenpairs/s "
        f"(capture {tcap:.2f} s)")
    return dt, tcap


def plain_path(M, torch, probs, stream, stride):
    """The plain path: one solve per matrix over every stride-th problem, same
    device-resident buffers.  Returns (eigenpairs/s, matrices solved)."""
    sample = list(range(0, len(probs), stride))
    singles = [M.prepare_batch([probs[i]])[0] for i in sample]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ppairs = 0
    for one in singles:
        rc = M.run_batch(one, stream=stream)
        assert rc == 0
        ppairs += one[0].m
    torch.cuda.synchronize()
    plain = ppairs / (time.perf_counter() - t0)
    log(f"plain: {plain:.4g} eigenpairs/s on {len(sample)} matrices")
    return plain, len(sample)


def roofline(prof, counted, prof_steps, traffic_key, dom=None):
    """Roofline of the dominant kernel (or of `dom`): its algorithmic flops per step
    (from the step run with the device work counters on) over its launches per step,
    against its average launch time in the timed steps.  Returns (kernel name,
    roofline dict)."""
    kern = prof["kernels"]
    if dom is None:
        dom = max(kern, key=lambda k: kern[k]["ms"])
    steps = counted["steps"]
    launches_per_step = kern[dom]["launches"] / prof_steps
    assert counted["kernels"][dom]["launches"] == launches_per_step, "launch count differs between steps"
    fams = KERNEL_STEPS.get(dom, [])
    flops = sum(steps.get(f, 0) * FLOPS_PER_STEP[f] for f in fams)
    per_launch_s = kern[dom]["ms"] / 1e3 / kern[dom]["launches"]
    flops_per_launch = flops / launches_per_step
    achieved = flops_per_launch / per_launch_s / 1e12 if per_launch_s > 0 else 0.0
    # the same kernel's ceiling when every SIMD issues division steps back to back
    dv = sum(DIVIDES_PER_STEP[f] * steps.get(f, 0) for f in fams)
    divide_ceiling = DIVIDE_STEPS_PER_S * flops / dv / 1e12 if dv else None
    # HBM bytes per launch measured by rocprofv3 on the same workload, when recorded
    traffic = None
    tfile = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            rec = json.load(open(tfile)).get(traffic_key, {})
            traffic = rec.get(dom) if isinstance(rec, dict) else None
        except Exception:
            traffic = None
    return dom, {"bound": "fp64_valu", "engine": "fp64 vector ALU (no MFMA on this path)",
                 "kernel": dom, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                 "divide_ceiling": divide_ceiling,
                 "frac_of_divide_ceiling": (achieved / divide_ceiling) if divide_ceiling else None,
                 "traffic": traffic, "flops_per_launch": flops_per_launch,
                 "ms_per_launch": per_launch_s * 1e3}


def serve_stream(M, count, rate, budget, inflight=3, batch=64, window=None):
    """The served stream through the server, open loop (Poisson arrivals at `rate`
    requests/s) or, with `window`, closed loop (at most `window` requests outstanding, a
    new one submitted as soon as one is taken back: the server's capacity).  A waiter
    thread takes the results back in submission order and keeps only latency and m, so
    that host memory holds the outstanding requests only."""
    import queue
    import threading
    from mr3smp_amd import workloads as WL
    reqs = WL.served_stream(count, rate if rate else 1.0)
    srv = M.Server(budget_bytes=budget, max_inflight=inflight, max_batch=batch)
    q = queue.Queue()
    got = []
    slots = threading.Semaphore(window) if window else None

    def waiter():
        while True:
            rid = q.get()
            if rid is None:
                return
            o = srv.wait(rid)
            got.append((o["t_done"] - o["t_submit"], o["m"], o["info"]))
            del o
            if slots:
                slots.release()

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    try:
        for r in reqs:
            if slots:
                slots.acquire()
            else:
                while time.perf_counter() - t0 < r["t"]:
                    time.sleep(min(0.002, max(0.0, r["t"] - (time.perf_counter() - t0))))
            kw = {k: r[k] for k in ("vl", "vu", "il", "iu") if k in r}
            mc = (r["iu"] - r["il"] + 1) if r["range"] == "I" else None
            q.put(srv.submit(r["jobz"], r["range"], r["D"], r["E"], mcols=mc, **kw))
    finally:
        q.put(None)
        th.join()
    dt = time.perf_counter() - t0
    srv.close()
    lat = np.array([g[0] for g in got])
    pairs = sum(g[1] for g in got)
    assert len(got) == count and all(g[2] == 0 for g in got)
    return dict(requests=count, rate_per_s=(count / dt) if window else rate,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                eigenpairs_per_s=pairs / dt, wall_s=dt,
                arrivals=f"closed loop, {window} outstanding" if window else "Poisson")


def serve_plain(M, count):
    """The first `count` requests of the served stream solved one after the other
    through the drop-in mrrr() (host arrays in and out), for a like-for-like plain rate."""
    from mr3smp_amd import workloads as WL
    reqs = WL.served_stream(count, 1.0)
    t0 = time.perf_counter()
    pairs = 0
    for r in reqs:
        kw = {k: r[k] for k in ("vl", "vu", "il", "iu") if k in r}
        mc = (r["iu"] - r["il"] + 1) if r["range"] == "I" else None
        o = M.mrrr(r["jobz"], r["range"], r["D"], r["E"], mcols=mc, **kw)
        assert o["info"] == 0
        pairs += o["m"]
    dt = time.perf_counter() - t0
    return dict(eigenpairs_per_s=pairs / dt, wall_s=dt, requests=count)


def cpu_baseline_subset(mats, stride, threads):
    """The compiled reference (oracle/_ref) on every stride-th matrix of the subset
    workload, one matrix at a time with PMR_NUM_THREADS=threads (its task-parallel
    route); the oracle instead, when the reference is not built, as `threads`
    concurrent callers (it is sequential)."""
    from concurrent.futures import ThreadPoolExecutor
    from mrrr_ctypes import ORACLE_SO, REF_SO, call_mrrr, load
    if os.path.exists(REF_SO):
        _, fn = load(REF_SO, "mrrr")
        kind = "reference"
    elif os.path.exists(ORACLE_SO):
        _, fn = load(ORACLE_SO, "oracle_mrrr")
        kind = "port"
    else:
        return None
    os.environ["PMR_NUM_THREADS"] = str(threads)
    sample = mats[::stride]

    def one(t):
        D, E, il, iu = t
        r = call_mrrr(fn, "V", "I", D, E, il=il, iu=iu, mcols=iu - il + 1)
        assert r["info"] == 0 and r["m"] == iu - il + 1
        return r["m"]

    t0 = time.perf_counter()
    if kind == "reference":
        pairs = sum(one(t) for t in sample)
        how = f"one matrix at a time with PMR_NUM_THREADS={threads}"
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            pairs = sum(ex.map(one, sample))
        how = f"{threads} concurrent callers"
    dt = time.perf_counter() - t0
    return dict(value=pairs / dt, unit="eigenpairs/s", cores=threads, kind=kind,
                sample=f"every {stride}th matrix of the subset workload ({len(sample)} matrices, "
                       f"{pairs} eigenpairs, jobz=V range=I), {how}, wall {dt:.2f} s")


def subset_main(args, torch, M, dev):
    """The subset workload: N(0,1) tridiagonals with n log-uniform in [4096, 50000],
    jobz V, range I over a seeded window of ceil(0.1 n) indices, all in one batch."""
    from mr3smp_amd import workloads as WL
    from mrrr_ctypes import resid_orth_device
    if args.autotune:
        sys.exit("--autotune searches on the mixed workload only")
    tuning_src = "library defaults"
    if args.tuning_file and os.path.exists(args.tuning_file):
        from mr3smp_amd import tune
        tune.apply(args.tuning_file)
        tuning_src = os.path.relpath(args.tuning_file, REPO)
    tuning = dict(M.get_tuning(), source=tuning_src)
    log("building workload")
    mats = WL.batched_range_i(args.subset_count)
    probs = [dict(D=torch.tensor(D, device=dev), E=torch.tensor(E, device=dev), jobz="V",
                  range="I", il=il, iu=iu, mcols=iu - il + 1) for D, E, il, iu in mats]
    pairs = sum(iu - il + 1 for _, _, il, iu in mats)
    stream = torch.cuda.Stream()
    arr, keep = M.prepare_batch(probs)
    dt, prof, counted = timed_steps(M, torch, arr, stream, args.steps, args.warmup, pairs)
    value = pairs * args.steps / dt
    plain, nplain = plain_path(M, torch, probs, stream, args.plain_stride)
    worst_r, worst_o = 0.0, 0.0
    for i in range(0, len(mats), 16):
        m = arr[i].m
        r, o = resid_orth_device(probs[i]["D"], probs[i]["E"], keep[i]["W"][:m], keep[i]["Z"][:m])
        worst_r, worst_o = max(worst_r, r), max(worst_o, o)
    dom, roof = roofline(prof, counted, args.steps, traffic_key="batched_subset")
    kern = prof["kernels"]
    z_bytes = 8.0 * sum(len(D) * (iu - il + 1) for D, _, il, iu in mats)
    norm_ms = kern.get("k_s_normalize", {}).get("ms", 0.0) / args.steps
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_subset(mats, args.cpu_stride or 32, min(16, os.cpu_count() or 1))
        log(f"cpu baseline: {cpu}")
    out = {
        "metric": "eigenpairs/sec batched vs plain on 1 GPU; p50/p99 latency; orth/resid <= ref",
        "value": value,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded), inputs and outputs resident in HBM",
        "config": {"workload": f"batched subset: {len(mats)} matrices, n log-uniform [4096,50000], "
                               f"D, E ~ N(0,1), jobz=V range=I over a window of ceil(0.1 n) indices",
                   "matrices": len(mats), "eigenpairs_per_step": pairs,
                   "z_bytes": z_bytes},
        "plain_value": plain,
        "batched_over_plain": value / plain,
        "plain_sample": f"every {args.plain_stride}th matrix ({nplain}), one solve per matrix",
        "accuracy_sample": {"max_resid": worst_r, "max_orth": worst_o, "matrices": len(range(0, len(mats), 16)),
                            "units": "n*eps*||T|| and n*eps"},
        "roofline": roof,
        "z_writeback": {"bytes_per_step": z_bytes,
                        "GBps_over_normalize_kernel": (z_bytes / (norm_ms / 1e3) / 1e9) if norm_ms else None,
                        "peak_GBps": HBM_PEAK_GBS},
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])},
        "launches_per_step": {k: v["launches"] / args.steps for k, v in kern.items()},
        "sweep_elements_per_step": dict(counted["steps"]),
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--plain-stride", type=int, default=8)
    ap.add_argument("--cpu-stride", type=int, default=None,
                    help="the CPU baseline runs on every k-th matrix (default 16 on the mixed "
                         "workload, 32 on the subset workload)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-window", type=int, default=48,
                    help="requests outstanding in the closed-loop capacity run")
    ap.add_argument("--serve-plain-requests", type=int, default=192,
                    help="requests of the stream solved one by one through mrrr() for the plain rate")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="arrival rate of the served stream in requests/s (default: 0.8 x the "
                         "capacity measured with every request arriving at once)")
    ap.add_argument("--serve-inflight", type=int, default=8,
                    help="server workers (concurrent batches)")
    ap.add_argument("--serve-batch", type=int, default=16, help="most requests per batch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured-graph form")
    ap.add_argument("--autotune", action="store_true",
                    help="search the launch-shape parameters on this workload before timing "
                         "(mr3smp_amd/tune.py) and write them to --tuning-file")
    ap.add_argument("--tuning-file", default=os.path.join(REPO, "mr3smp_amd", "tuning_mi355x.json"),
                    help="launch-shape parameters applied before timing, if the file exists")
    ap.add_argument("--workload", choices=["mixed", "subset"], default="mixed",
                    help="mixed: the batched mixed workload (the headline line); subset: 256 "
                         "N(0,1) tridiagonals, n log-uniform in [4096, 50000], jobz V, range I "
                         "over a 10%% window, in one batch")
    ap.add_argument("--subset-count", type=int, default=256)
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (five times): the "
                         "known-size kernel that scales the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    # HIP's hardware queues per process (GPU_MAX_HW_QUEUES) are left as the environment
    # has them (HIP's default is 4); the solver picks its number of side streams from it
    import torch
    import mr3smp_amd as M
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    if args.copy_calib:
        nel = args.copy_calib * (1 << 20) // 8
        src = torch.rand(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_each_way": nel * 8, "copies": 5}))
        return
    if args.workload == "subset":
        subset_main(args, torch, M, dev)
        return
    log("building workload")
    mats, probs = build_workload(torch, dev, args.count)
    pairs = sum(len(D) for D, _, _ in mats)
    stream = torch.cuda.Stream()
    # the problem descriptors (pointers into HBM and scalars) are built once; a step
    # is one mr3_solve_batch call over them
    arr, keep = M.prepare_batch(probs)
    tuning_src = "library defaults"
    if args.autotune:
        from mr3smp_amd import tune
        best, bt, tried = tune.autotune(arr, stream=stream, log=log)
        tune.save(best, args.tuning_file,
                  note=f"coordinate descent on the batched mixed workload ({len(tried)} settings, "
                       f"best {bt * 1e3:.1f} ms per solve)")
        tuning_src = f"autotuned in this run ({len(tried)} settings)"
    elif args.tuning_file and os.path.exists(args.tuning_file):
        from mr3smp_amd import tune
        tune.apply(args.tuning_file)
        tuning_src = os.path.relpath(args.tuning_file, REPO)
    tuning = dict(M.get_tuning(), source=tuning_src)
    dt, prof, counted = timed_steps(M, torch, arr, stream, args.steps, args.warmup, pairs)
    eager_value = pairs * args.steps / dt
    prof_steps = args.steps
    # the same steps replayed from one captured graph; the headline is the faster form
    if args.no_graph:
        gdt, gcap = float("inf"), None
    else:
        gdt, gcap = graph_steps(M, torch, arr, stream, args.steps, pairs)
    graph_value = pairs * args.steps / gdt
    launch = "graph" if graph_value > eager_value else "eager"
    value = max(graph_value, eager_value)
    step_s = gdt if launch == "graph" else dt
    plain, nplain = plain_path(M, torch, probs, stream, args.plain_stride)

    worst_r, worst_o = accuracy_sample(mats, keep, arr, 64)

    dom, roof = roofline(prof, counted, prof_steps, traffic_key="batched_mixed")
    # the task-parallel pipeline's own dominant kernel beside it (k_small, the small
    # matrices' route, runs concurrently with it)
    roof_pipe = None
    if dom == "k_small" and "k_s_iter" in prof["kernels"]:
        roof_pipe = roofline(prof, counted, prof_steps, traffic_key="batched_mixed", dom="k_s_iter")[1]
    kern = prof["kernels"]
    steps = counted["steps"]
    total_flops = sum(steps.get(f, 0) * FLOPS_PER_STEP[f] for f in FLOPS_PER_STEP) * prof_steps
    gpu_ms = sum(v["ms"] for v in kern.values())
    z_bytes = 8.0 * sum(len(D) ** 2 for D, _, _ in mats) * prof_steps
    # Z is written once in full by the normalisation pass (support scaled, rest zero)
    norm_ms = kern.get("k_s_normalize", {}).get("ms", 0.0)

    serve = None
    if not args.no_serve:
        try:
            # the server's device-memory budget: what this process has not taken yet (the
            # batched workload's buffers and the library's cached workspaces stay
            # allocated), with room left for the runtime's own scratch allocations
            free_b, _ = torch.cuda.mem_get_info()
            budget = int(min(200 << 30, 0.6 * free_b))
            cap = None
            rate = args.serve_rate
            if rate is None:
                # capacity: the same requests, a closed loop that keeps the server busy
                cap = serve_stream(M, args.serve_requests, None, budget,
                                   args.serve_inflight, args.serve_batch, window=args.serve_window)
                rate = 0.8 * args.serve_requests / cap["wall_s"]
                log(f"serve capacity: {cap}")
            serve = serve_stream(M, args.serve_requests, rate, budget, args.serve_inflight,
                                 args.serve_batch)
            serve["server"] = {"max_inflight": args.serve_inflight, "max_batch": args.serve_batch,
                               "budget_GiB": round(budget / 2**30, 1)}
            if cap is not None:
                serve["capacity"] = {"requests_per_s": args.serve_requests / cap["wall_s"],
                                     "eigenpairs_per_s": cap["eigenpairs_per_s"],
                                     "p50_ms": cap["p50_ms"], "p99_ms": cap["p99_ms"]}
                serve["rate_rule"] = "0.8 x capacity"
                plain_s = serve_plain(M, args.serve_plain_requests)
                serve["plain_first_requests"] = plain_s
                serve["capacity_over_plain"] = cap["eigenpairs_per_s"] / plain_s["eigenpairs_per_s"]
            log(f"serve: {serve}")
        except Exception as e:  # reported, not fatal for the headline number
            serve = {"error": repr(e)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mats, args.cpu_stride or 16, min(16, os.cpu_count() or 1))
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "eigenpairs/sec batched vs plain on 1 GPU; p50/p99 latency; orth/resid <= ref",
        "value": value,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded), inputs and outputs resident in HBM",
        "config": {"workload": f"batched mixed tridiagonals: {args.count} matrices, n log-uniform "
                               f"[32,2048], 50% U(-1,1) / 25% perturbed 1-2-1 / 25% glued W21+ "
                               f"(glue 1e-10), jobz=V range=A",
                   "matrices": args.count, "eigenpairs_per_step": pairs, "launch": launch},
        "eager": {"eigenpairs_per_s": eager_value, "ms_per_step": dt / args.steps * 1e3,
                  "how": "mr3_solve_batch per step"},
        "graph": None if args.no_graph else {
            "eigenpairs_per_s": graph_value, "ms_per_step": gdt / args.steps * 1e3,
            "capture_s": gcap, "how": "one captured hipGraph replayed per step (mr3_graph_launch)"},
        "plain_value": plain,
        "batched_over_plain": value / plain,
        "plain_sample": f"every {args.plain_stride}th matrix, one solve per matrix",
        "serve": serve,
        "accuracy_sample": {"max_resid": worst_r, "max_orth": worst_o,
                            "units": "n*eps*||T|| and n*eps"},
        "roofline": roof,
        "roofline_pipeline": roof_pipe,
        "z_writeback": {"bytes_per_step": z_bytes / prof_steps,
                        "GBps_over_normalize_kernel": (z_bytes / (norm_ms / 1e3) / 1e9) if norm_ms else None,
                        "peak_GBps": HBM_PEAK_GBS},
        "kernels_ms_per_step": {k: v["ms"] / prof_steps for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])},
        "gpu_fp64_tflops_all_kernels": total_flops / (gpu_ms / 1e3) / 1e12 if gpu_ms else None,
        "launches_per_step": {k: v["launches"] / prof_steps for k, v in kern.items()},
        "sweep_elements_per_step": dict(steps),
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
