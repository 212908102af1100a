"""Benchmark of the gfx950 xtax annealer (one GPU).

Default workload (`--workload batch`, the metric's configuration): 1024 planted
instances, n ~ U{8..48}, A = U^T D U with D = diag(+-U{1..100}) and U a product
of 3n random +-1 shears (xtax_amd.workloads.batch_config), 32 tempering
replicas each on a geometric ladder T in [14, 18], greedy fraction 0.5,
target fraction 1.0,
exchange every 2000 moves, all instances resident in one captured launch
sequence. An instance retires when a replica reaches a diagonal or after
`--cap-epochs` epochs. One step = one pass over the whole batch. Every solved
result is re-verified exactly on the host (X^T A0 X = A, A diagonal,
|det X| = 1) and only verified ones count. The launch configuration of each
order bucket is autotuned once before the timed steps (`--no-tune`: the rule).
`value` counts each step from the moment its inputs are resident in HBM (the
host-to-device upload is excluded and reported under `with_upload`).

Prints one JSON line: exact-diagonalized matrices per second, ms_per_step,
solve-latency percentiles (overall and per order bucket, with the solved
fraction), shear moves per second, the plain form's matrices/s on the first
`--plain-count` instances with the same settings (one instance per launch) and
the batch/plain ratio, the roofline of the dominant kernel (k_anneal_lane) and a
CPU baseline: the reference's own xtax binary (oracle/_ref/xtax_ref, compiled
from the reference sources) on bounded samples of the same instances on this
host's cores, one instance at a time on every core and as concurrent
single-thread runs, each the median of three runs.

The ladders and the target fraction are per workload (--t-init, --t-min and
--target-fraction override them), chosen by solve rate on the workloads' own
instances: T in [14, 18] for the batch (D up to 100; [5, 200] solved 1.75x
fewer), [36, 50] for serving (D up to 200), and every proposal targeted
(--target-fraction 1: 1.4x the solves of the reference default 0.5). The CPU
baseline runs the reference with the same settings.

Other configurations (`--workload ...`), same line format:
  plain    the batch instances one per launch (xtx_solve_plain), first --count
  serve    Poisson stream, n = round(exp(U[ln 6, ln 64])), D = +-U{1..200},
           64 replicas per request, continuous batching (xtx_serve_*)
  large    one planted n = 80 instance (D = +-U{1..1000}, 4n shears),
           4096 replicas, greedy 0.7, T in [50, 2000]; --count K: time to an
           exact diagonal on K instances, each capped at --seconds;
           --x-in-lds: X^T in LDS instead of HBM
  lattice  24x24 lattice Gram (entries U[-9, 9]), 256 replicas, greedy 0.5,
           sweep 0.05
  random   n = 10 random symmetric (U[-20, 20]), 8 replicas, T in [0.5, 50],
           one seed per launch; "batched" = the same seeds as one batch
  dual     xdual's primal/dual objective (xtx_dual_solve) on the lattice
           workload's Gram, 4096 replicas, greedy 0.5; CPU baseline the
           reference's xdual binary (oracle/_ref/xdual_ref)
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from xtax_amd import DualParams, Engine, EngineParams, workloads  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "xtax_ref")
REF_DUAL_BIN = os.path.join(ROOT, "oracle", "_ref", "xdual_ref")
# 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz: int32 VALU lane-ops/s of an MI355X
MI355X_INT_LANE_OPS = 256 * 64 * 2.4e9
METRIC = "exact-diagonalized matrices/sec + p50/p99 solve latency; shear moves/sec"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, float), q)) if len(xs) else None


def base_line(args, value, unit, el, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": el * 1e3 / max(1, args.steps),
            "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": "int64", "data": "synthetic", "config": config}


def measured_traffic(profile, per="epoch"):
    """HBM bytes of one epoch (or one launch) of the dominant kernel from its
    committed rocprof summary (FETCH_SIZE and WRITE_SIZE passes scaled by the
    copy-kernel calibration), or None."""
    import re
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            m = re.search(r"traffic: ([0-9.]+e[0-9]+) bytes per " + per, f.read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def roofline(kernel_ms, moves, mean_n, greedy, kernel="k_anneal"):
    """Algorithmic integer work of the annealing kernel over its measured time.

    Per proposal the method evaluates one Add (6n int64 ops over the n
    columns: multiply, add, bound check, two absolute values, accumulate) and,
    with probability `greedy`, up to five candidate shears (5 x 6n) plus the
    weighted-median search (9 weighted counts, 3n each). An int64 op is counted
    as 2 int32 lane-ops. Peak: int32 VALU lane-ops of the MI355X."""
    if not kernel_ms:
        return None
    ops_per_move = 6 * mean_n * (1 + greedy * 5) + greedy * 9 * 3 * mean_n
    achieved = moves * ops_per_move * 2 / (kernel_ms / 1e3) / 1e12
    peak = MI355X_INT_LANE_OPS / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
            "frac": achieved / peak, "traffic": None,
            "kernel": kernel, "kernel_ms": kernel_ms, "moves": moves,
            "note": "integer VALU issue bound: no dense contraction for MFMA, and HBM is "
                    "touched only for X^T rows and epoch boundaries"}


def run_ref(A, args_list, timeout, binary=REF_BIN):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "A.csv")
        with open(path, "w") as f:
            for row in A:
                f.write(",".join(str(int(v)) for v in row) + "\n")
        cp = subprocess.run([binary, "-A", path] + args_list, cwd=td, capture_output=True,
                            text=True, timeout=timeout)
    moves = 0
    for line in cp.stdout.splitlines():
        if line.startswith("done moves="):
            moves = int(line.split()[1].split("=")[1])
    best = None
    for line in cp.stdout.splitlines():
        if line.startswith("Final best score:"):
            best = float(line.split(":")[1])
    return "matrix is diagonal" in cp.stdout, moves, best


def host_cores():
    """Cores this process may run on (the box's CPU share, not the machine's)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return None


def cpu_concurrent(mats, extra, budget_s, cap_s, workers):
    """Mode (ii): `workers` concurrent single-thread reference runs (-t 1,
    cooling schedule, as the reference runs one worker) over the instances in
    order, each capped at cap_s seconds, until budget_s is spent."""
    import concurrent.futures as cf
    t0 = time.perf_counter()
    nxt = 0
    solved = moves = done = 0

    def one(q):
        return run_ref(mats[q], ["-t", "1", "--seed", str(q + 1), "--max-seconds", str(cap_s)]
                       + extra, timeout=cap_s + 30)

    with cf.ThreadPoolExecutor(workers) as ex:
        running = set()
        while nxt < len(mats) and len(running) < workers:
            running.add(ex.submit(one, nxt))
            nxt += 1
        while running:
            fin, running = cf.wait(running, return_when=cf.FIRST_COMPLETED)
            for f in fin:
                ok, mv, _ = f.result()
                done += 1
                moves += mv
                solved += ok
            while (nxt < len(mats) and len(running) < workers
                   and time.perf_counter() - t0 < budget_s):
                running.add(ex.submit(one, nxt))
                nxt += 1
    el = time.perf_counter() - t0
    return {"value": solved / el, "solved": solved, "instances": done, "seconds": el,
            "moves_per_sec": moves / el}


def cpu_baseline_modes(mats, extra, budget_s, cap_s, cap_single_s, repeats=3):
    """The reference binary on this host, both ways BASELINE.md asks, each the
    median of `repeats` runs: (i) one instance at a time with every core
    (-t cores), (ii) `cores` concurrent single-thread instances."""
    if not os.path.exists(REF_BIN):
        return None
    cores = host_cores()
    runs_i = [cpu_sample(mats, extra, budget_s, cap_s, cores) for _ in range(repeats)]
    runs_ii = [cpu_concurrent(mats, extra, budget_s, cap_single_s, cores)
               for _ in range(repeats)]
    med_i = sorted(runs_i, key=lambda r: r["value"])[repeats // 2]
    med_ii = sorted(runs_ii, key=lambda r: r["value"])[repeats // 2]
    best = med_i if med_i["value"] >= med_ii["value"] else med_ii
    return {"value": best["value"], "unit": "matrices/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": (f"median of {repeats} runs of {budget_s:.0f} s each way over the batch "
                       f"instances in order: (i) xtax_ref -t {cores} {' '.join(extra)} "
                       f"--max-seconds {cap_s} one at a time, {med_i['value']:.3f} matrices/s "
                       f"({med_i['solved']} of {med_i['instances']}); (ii) {cores} concurrent "
                       f"xtax_ref -t 1 --max-seconds {cap_single_s}, {med_ii['value']:.3f} "
                       f"matrices/s ({med_ii['solved']} of {med_ii['instances']}); value = "
                       f"the faster mode"),
            "mode_i": {"runs": [r["value"] for r in runs_i], "median": med_i["value"],
                       "moves_per_sec": med_i["moves_per_sec"],
                       "p50_latency_s": med_i["p50_latency_s"]},
            "mode_ii": {"runs": [r["value"] for r in runs_ii], "median": med_ii["value"],
                        "moves_per_sec": med_ii["moves_per_sec"]}}


def cpu_sample(mats, extra, budget_s, cap_s, threads=None):
    """The reference binary on the first instances, one at a time with every
    core, each capped at cap_s seconds, until budget_s is spent."""
    if not os.path.exists(REF_BIN):
        return None
    cores = threads or host_cores()
    t0 = time.perf_counter()
    solved = moves = done = 0
    lat = []
    for q, A in enumerate(mats):
        if time.perf_counter() - t0 > budget_s:
            break
        t1 = time.perf_counter()
        ok, mv, _ = run_ref(A, ["-t", str(cores), "--seed", str(q + 1), "--max-seconds",
                                str(cap_s)] + extra, timeout=cap_s + 30)
        done += 1
        moves += mv
        if ok:
            solved += 1
            lat.append(time.perf_counter() - t1)
    el = time.perf_counter() - t0
    return {"value": solved / el, "unit": "matrices/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} instances one at a time: xtax_ref -t {cores} "
                      f"{' '.join(extra)} --max-seconds {cap_s}; {solved} solved in {el:.1f} s",
            "moves_per_sec": moves / el, "p50_latency_s": pct(lat, 50), "solved": solved,
            "instances": done}


def warm(eng):
    tri = (2 * np.eye(4, dtype=np.int64) + np.eye(4, k=1, dtype=np.int64)
           + np.eye(4, k=-1, dtype=np.int64))
    eng.run_annealer(tri, None, EngineParams(replicas=2, max_epochs=1))


# --------------------------------------------------------------------- batch
def bench_batch(args, eng):
    mats = workloads.batch_config(count=args.count, seed=7)
    seeds = [args.seed * 1_000_003 + q for q in range(len(mats))]
    mean_n = float(np.mean([a.shape[0] for a in mats]))
    p = EngineParams(replicas=args.replicas, t_init=args.t_init, t_min=args.t_min,
                     greedy_fraction=0.5, target_fraction=args.target_fraction,
                     exchange_interval=args.exchange_interval,
                     max_epochs=args.cap_epochs, seed=args.seed, verify=True,
                     lookahead=args.lookahead, persistent=not args.no_persistent)
    tune = None
    if not args.no_tune:
        # launch configuration per order bucket, measured once per device
        # before the timed steps (like a compile); results do not depend on it
        t1 = time.perf_counter()
        rows = eng.autotune(mats, p, moves=512, seeds=seeds)
        tune = {"seconds": time.perf_counter() - t1,
                "chosen": [{k: r[k] for k in ("order", "lanes", "cols", "triangular",
                                              "waves_per_block", "lookahead", "moves_per_sec")}
                           for r in rows if r["tuned"]]}
        log(f"autotune {tune['seconds']:.2f}s: " + ", ".join(
            f"{c['order']}:{c['lanes']}x{c['cols']}{'t' if c['triangular'] else 's'}"
            f"w{c['waves_per_block']}" + (f"k{c['lookahead']}" if c["lookahead"] > 1 else "")
            for c in tune["chosen"]))
    for w in range(args.warmup):
        eng.batch_solve(mats, p, seeds=seeds)
    solved = verified = moves = 0
    lat = []
    kernel_ms = 0.0
    lane_instances = 0
    persistent = 0
    resident_s = 0.0  # each step less its upload of the host inputs
    buckets = {}  # n bucket -> [instances, solved, latencies]
    t0 = time.perf_counter()
    for k in range(args.steps):
        res, bst = eng.batch_solve(mats, p, seeds=seeds)
        for q, (A, X, d) in enumerate(res):
            b = buckets.setdefault(bucket_of(mats[q].shape[0]), [0, 0, []])
            b[0] += 1
            if d["solved"]:
                solved += 1
                if d["verify_status"] == 0:
                    verified += 1
                    lat.append(d["retire_s"] - d["submit_s"])
                    b[1] += 1
                    b[2].append(d["retire_s"] - d["submit_s"])
        moves += bst["moves"]
        kernel_ms += bst["kernel_ms"]
        lane_instances += bst["lane_instances"]
        persistent += bst["persistent"]
        resident_s += bst["seconds"] - bst["upload_seconds"]
        log(f"step {k}: solved {bst['solved']}/{bst['instances']} verified {bst['verified']} "
            f"epochs {bst['epochs']} buckets {bst['buckets']} "
            f"{'persistent ' if bst['persistent'] else ''}{bst['seconds']:.2f}s "
            f"(upload {bst['upload_seconds']:.3f}s) kernel {bst['kernel_ms']:.0f}ms "
            f"moves/s {bst['moves'] / bst['seconds']:.3e}")
    el = time.perf_counter() - t0
    cfg = {"workload": f"batch: {args.count} planted instances n~U{{8..48}}, D=+-U{{1..100}}, "
                       f"3n shears, {args.replicas} replicas, T in [{args.t_min},{args.t_init}], "
                       f"greedy 0.5, target {args.target_fraction}, exchange "
                       f"{args.exchange_interval}, cap {args.cap_epochs} epochs",
           "instances": args.count, "replicas": args.replicas,
           "cap_moves_per_replica": args.cap_epochs * args.exchange_interval,
           "launch": "one persistent launch per step" if persistent == args.steps
                     else "captured graphs of epochs"}
    # value: inputs resident in HBM when the clock starts (each step's own
    # upload excluded); the rate with the host-to-device upload is reported
    # beside it
    out = base_line(args, verified / resident_s, "matrices/s", resident_s, cfg)
    out.update({"solved": solved, "verified": verified,
                "solved_fraction": solved / (args.count * args.steps),
                "with_upload": {"value": verified / el, "seconds": el},
                "autotune": tune,
                "p50_latency_s": pct(lat, 50), "p95_latency_s": pct(lat, 95),
                "p99_latency_s": pct(lat, 99), "moves_per_sec": moves / resident_s,
                "kernel_moves_per_sec": moves / (kernel_ms / 1e3) if kernel_ms else None,
                "roofline": roofline(kernel_ms, moves, mean_n, 0.5, kernel_name(
                    lane_instances, args.count * args.steps)), "cpu_baseline": None})
    if out["roofline"] and persistent == args.steps and lane_instances == args.count * args.steps:
        # the whole step is one launch of the persistent kernel; its HBM bytes
        # per launch measured by rocprof on this workload
        out["roofline"]["kernel"] = "k_lane_persist"
        out["roofline"]["traffic"] = measured_traffic("k_lane_persist_batch1024.md", "launch")
        out["roofline"]["traffic_unit"] = "bytes per launch (profiles/k_lane_persist_batch1024.md)"
    elif out["roofline"] and out["roofline"]["kernel"] == "k_anneal_lane":
        # HBM bytes per epoch (the three launches of the lane kernel's order
        # pairs) measured by rocprof on this workload
        out["roofline"]["traffic"] = measured_traffic("k_anneal_lane_batch1024.md")
        out["roofline"]["traffic_unit"] = "bytes per epoch (profiles/k_anneal_lane_batch1024.md)"
    out["by_order"] = {
        name: {"instances": b[0], "solved_fraction": b[1] / b[0],
               "p50_latency_s": pct(b[2], 50), "p99_latency_s": pct(b[2], 99)}
        for name, b in sorted(buckets.items(), key=lambda kv: int(kv[0].split("-")[0]))}
    if args.plain_count > 0:
        # the plain form (one instance per launch, xtx_solve_plain) on the first
        # instances of the same batch with the same settings and seeds
        out["plain_form"] = plain_leg(eng, mats[:args.plain_count], seeds, p)
        out["batch_over_plain"] = out["value"] / max(out["plain_form"]["value"], 1e-12)
        log(f"plain form: {out['plain_form']['value']:.3f} matrices/s "
            f"({out['plain_form']['solved']} of {args.plain_count}); batch/plain "
            f"{out['batch_over_plain']:.1f}x")
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline_modes(
            mats, ["--t-init", str(args.t_init), "--t-min", str(args.t_min),
                   "--greedy-fraction", "0.5", "--target-fraction", str(args.target_fraction)],
            args.cpu_seconds, args.cpu_cap, args.cpu_cap_single)
    return out


def kernel_name(lane_instances, instances):
    """The kernel family the batch ran: the lane kernel (one lane per
    proposal) when the replicas fill the device, else the lane-group one."""
    if lane_instances == instances:
        return "k_anneal_lane"
    if lane_instances == 0:
        return "k_anneal"
    return f"k_anneal_lane ({lane_instances} of {instances} instances) + k_anneal"


def bucket_of(n):
    lo = (n // 8) * 8
    return f"{lo}-{lo + 7}"


def plain_leg(eng, mats, seeds, p):
    """matrices/s of the plain form: each instance alone, one per launch."""
    solved = moves = 0
    lat = []
    t0 = time.perf_counter()
    for q, A in enumerate(mats):
        pq = EngineParams(**{**p.__dict__, "seed": seeds[q]})
        t1 = time.perf_counter()
        r = eng.run_annealer(A, None, pq)
        moves += r.stats["moves"]
        if r.stats["solved"] and r.stats["verify_status"] == 0:
            solved += 1
            lat.append(time.perf_counter() - t1)
    el = time.perf_counter() - t0
    return {"value": solved / el, "unit": "matrices/s", "instances": len(mats),
            "solved": solved, "seconds": el, "moves_per_sec": moves / el,
            "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
            "form": "xtx_solve_plain per instance (inputs uploaded per call), same settings "
                    "and seeds as the batch"}


# --------------------------------------------------------------------- plain
def bench_plain(args, eng):
    mats = workloads.batch_config(count=1024, seed=7)[:args.count]
    seeds = [args.seed * 1_000_003 + q for q in range(len(mats))]
    solved = moves = 0
    lat = []
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for q, A in enumerate(mats):
        p = EngineParams(replicas=args.replicas, t_init=args.t_init, t_min=args.t_min,
                         greedy_fraction=0.5, target_fraction=args.target_fraction,
                         exchange_interval=args.exchange_interval,
                         max_epochs=args.cap_epochs, seed=seeds[q], verify=True,
                         lookahead=args.lookahead)
        t1 = time.perf_counter()
        r = eng.run_annealer(A, None, p)
        moves += r.stats["moves"]
        kernel_ms += r.stats["kernel_ms"]
        if r.stats["solved"] and r.stats["verify_status"] == 0:
            solved += 1
            lat.append(time.perf_counter() - t1)
    el = time.perf_counter() - t0
    mean_n = float(np.mean([a.shape[0] for a in mats]))
    cfg = {"workload": f"plain: the first {len(mats)} batch instances one per launch, "
                       f"{args.replicas} replicas, T in [{args.t_min},{args.t_init}], target "
                       f"{args.target_fraction}, cap {args.cap_epochs} epochs"}
    out = base_line(args, solved / el, "matrices/s", el, cfg)
    out.update({"solved": solved, "instances": len(mats), "p50_latency_s": pct(lat, 50),
                "p99_latency_s": pct(lat, 99), "moves_per_sec": moves / el,
                # one instance per launch: its 32 replicas are too few for the
                # lane kernel, so the lookahead lane-group kernel runs
                "roofline": roofline(kernel_ms, moves, mean_n, 0.5, "k_anneal"),
                "cpu_baseline": None})
    return out


# --------------------------------------------------------------------- serve
def bench_serve(args, eng):
    count = args.count
    reqs = workloads.serve_config(count=count, dmax=200, seed=11, rate=args.rate)
    p = EngineParams(replicas=64, t_init=args.t_init, t_min=args.t_min, greedy_fraction=0.5,
                     target_fraction=args.target_fraction,
                     exchange_interval=args.exchange_interval, max_epochs=args.cap_epochs,
                     seed=args.seed, verify=True, lookahead=args.lookahead)
    eng.serve_open(p)
    lat_solved, lat_all = [], []
    solved = verified = moves = 0
    done = 0
    nxt = 0
    t0 = time.perf_counter()
    arrivals = {}
    next_report = 20.0
    while done < count:
        now = time.perf_counter() - t0
        if now >= next_report:
            log(f"serve: {now:.0f}s, {done}/{count} retired, {solved} solved")
            next_report += 20.0
        while nxt < count and reqs[nxt][0] <= now:
            eng.serve_submit(reqs[nxt][1], nxt, args.seed * 1_000_003 + nxt)
            arrivals[nxt] = t0 + reqs[nxt][0]
            nxt += 1
        resident, queued = eng.serve_step()
        for A, X, d in eng.serve_poll():
            t_ret = time.perf_counter()
            lat = t_ret - arrivals.pop(d["tag"])
            lat_all.append(lat)
            moves += d["moves"]
            done += 1
            if d["solved"]:
                solved += 1
                if d["verify_status"] == 0:
                    verified += 1
                    lat_solved.append(lat)
        if resident == 0 and queued == 0 and nxt < count:
            time.sleep(max(0.0, min(0.01, reqs[nxt][0] - (time.perf_counter() - t0))))
    el = time.perf_counter() - t0
    eng.serve_close()
    cfg = {"workload": f"serve: {count} planted requests, Poisson {args.rate}/s, "
                       f"n=round(exp(U[ln6,ln64])), D=+-U{{1..200}}, 64 replicas, "
                       f"T in [{args.t_min},{args.t_init}], target {args.target_fraction}, "
                       f"cap {args.cap_epochs} epochs"}
    out = base_line(args, verified / el, "matrices/s", el, cfg)
    out.update({"solved": solved, "verified": verified, "requests": count,
                "p50_latency_s": pct(lat_solved, 50), "p95_latency_s": pct(lat_solved, 95),
                "p99_latency_s": pct(lat_solved, 99),
                "p50_latency_all_s": pct(lat_all, 50), "p99_latency_all_s": pct(lat_all, 99),
                "moves_per_sec": moves / el, "cpu_baseline": None})
    return out


# ------------------------------------------------------- single-instance runs
def bench_single(args, eng, name, A0, params, cpu_extra):
    t0 = time.perf_counter()
    r = eng.run_annealer(A0, None, params)
    el = time.perf_counter() - t0
    n = A0.shape[0]
    cfg = {"workload": name, "n": n, "replicas": params.replicas}
    out = base_line(args, r.stats["moves"] / el, "moves/s", el, cfg)
    out.update({"solved": r.stats["solved"], "best_score": r.stats["best_score"],
                "start_score": int(2 * np.abs(A0).sum() - np.abs(np.diag(A0)).sum()),
                "epochs": r.stats["epochs"], "verify_status": r.stats["verify_status"],
                "roofline": roofline(r.stats["kernel_ms"], r.stats["moves"], n,
                                     params.greedy_fraction), "cpu_baseline": None})
    if not args.no_cpu and os.path.exists(REF_BIN):
        cores = os.cpu_count() or 1
        ok, mv, best = run_ref(A0, ["-t", str(cores), "--seed", "1", "--max-seconds",
                                    str(args.cpu_cap)] + cpu_extra, timeout=args.cpu_cap + 30)
        out["cpu_baseline"] = {"value": mv / args.cpu_cap, "unit": "moves/s", "cores": cores,
                               "kind": "reference", "best_score": best, "solved": ok,
                               "sample": f"xtax_ref -t {cores} {' '.join(cpu_extra)} "
                                         f"--max-seconds {args.cpu_cap}"}
    return out


def bench_large(args, eng):
    A0 = workloads.planted(80, 1000, 320, seed=5)[0]
    p = EngineParams(replicas=4096, greedy_fraction=0.7, t_init=args.t_init,
                     t_min=args.t_min, exchange_interval=args.exchange_interval,
                     max_seconds=args.seconds, seed=args.seed, verify=True,
                     lookahead=args.lookahead)
    name = "large: planted n=80, D=+-U{1..1000}, 4n shears, 4096 replicas, greedy 0.7"
    if args.x_in_lds:
        eng.launch_set(80, 64, 2, 1, 0, 4, x_in_lds=1)
        name += ", X^T in LDS"
    extra = ["--greedy-fraction", "0.7", "--t-init", str(args.t_init), "--t-min", str(args.t_min)]
    if (args.count or 1) <= 1:
        return bench_single(args, eng, name, A0, p, extra)
    return large_to_diagonal(args, eng, name, p)


def large_to_diagonal(args, eng, name, p):
    """Time to an exact diagonal on `--count` planted n = 80 instances
    (planted seeds 5, 6, ...), each run until a replica reaches a diagonal or
    `--seconds` pass; every solved result verified exactly."""
    runs = []
    moves = 0
    el_total = 0.0
    for q in range(args.count):
        A0 = workloads.planted(80, 1000, 320, seed=5 + q)[0]
        t0 = time.perf_counter()
        r = eng.run_annealer(A0, None, p)
        el = time.perf_counter() - t0
        el_total += el
        moves += r.stats["moves"]
        ok = bool(r.stats["solved"]) and r.stats["verify_status"] == 0
        start = int(2 * np.abs(A0).sum() - np.abs(np.diag(A0)).sum())
        runs.append({"planted_seed": 5 + q, "solved": ok, "seconds": el,
                     "best_score": r.stats["best_score"], "start_score": start,
                     "moves": r.stats["moves"]})
        log(f"large instance {q}: solved {ok} in {el:.1f} s, best {r.stats['best_score']} "
            f"from {start}, {r.stats['moves'] / el:.3e} moves/s")
    solved = [x for x in runs if x["solved"]]
    cfg = {"workload": name + f", {args.count} instances, cap {args.seconds:.0f} s each",
           "n": 80, "replicas": p.replicas, "instances": args.count}
    out = base_line(args, moves / el_total, "moves/s", el_total, cfg)
    out.update({"solved": len(solved), "solved_fraction": len(solved) / args.count,
                "time_to_diagonal_s": [x["seconds"] for x in solved],
                "p50_time_to_diagonal_s": pct([x["seconds"] for x in solved], 50),
                "cap_seconds": args.seconds, "runs": runs, "cpu_baseline": None})
    return out


def bench_lattice(args, eng):
    L, G = workloads.lattice(24, 24, seed=2)
    p = EngineParams(replicas=256, greedy_fraction=0.5, sweep_fraction=0.05,
                     exchange_interval=args.exchange_interval, max_seconds=args.seconds,
                     seed=args.seed, verify=True, lookahead=args.lookahead)
    return bench_single(args, eng, "lattice: 24x24 basis U[-9,9], Gram, 256 replicas, greedy "
                        "0.5, sweep 0.05", G, p,
                        ["--greedy-fraction", "0.5", "--sweep-fraction", "0.05"])


def bench_random(args, eng):
    seeds = range(1, args.count + 1)
    mats = [workloads.random_symmetric(10, -20, 20, seed=s) for s in seeds]

    def params(seed):
        return EngineParams(replicas=8, t_init=50.0, t_min=0.5, greedy_fraction=0.5,
                            exchange_interval=args.exchange_interval,
                            max_epochs=args.cap_epochs, seed=seed, verify=True,
                            lookahead=args.lookahead)

    solved = moves = 0
    plain = []
    t0 = time.perf_counter()
    for s, A0 in zip(seeds, mats):
        r = eng.run_annealer(A0, None, params(s))
        plain.append((r.A, r.X))
        solved += r.stats["solved"]
        moves += r.stats["moves"]
    el = time.perf_counter() - t0
    cfg = {"workload": f"random: {args.count} seeds, n=10 symmetric U[-20,20], 8 replicas, "
                       f"T in [0.5,50], greedy 0.5, cap {args.cap_epochs} epochs"}
    out = base_line(args, moves / el, "moves/s", el, cfg)
    out.update({"solved": solved, "instances": args.count, "cpu_baseline": None})
    # The same seeds as one batch: the same trajectories run concurrently
    # (each instance's A and X equal its plain run's, checked here).
    res, bst = eng.batch_solve(mats, params(1), seeds=list(seeds))
    same = all(np.array_equal(res[q][0], plain[q][0]) and np.array_equal(res[q][1], plain[q][1])
               for q in range(len(mats)))
    resident = bst["seconds"] - bst["upload_seconds"]
    out["batched"] = {"value": bst["moves"] / resident, "unit": "moves/s",
                      "seconds": resident, "solved": bst["solved"], "equal_to_plain": same}
    log(f"batched: {bst['moves'] / resident:.3e} moves/s, solved {bst['solved']}, "
        f"results equal to the plain runs: {same}")
    if not args.no_cpu and os.path.exists(REF_BIN):
        # the reference on the first seeds, 8 threads = 8 replicas as here
        extra = ["-t", "8", "--t-init", "50", "--t-min", "0.5", "--greedy-fraction", "0.5"]
        cmv = csolved = 0
        t1 = time.perf_counter()
        for s in list(seeds)[:4]:
            A0 = workloads.random_symmetric(10, -20, 20, seed=s)
            ok, mv, _ = run_ref(A0, extra + ["--seed", str(s), "--max-seconds", str(args.cpu_cap)],
                                timeout=args.cpu_cap + 30)
            cmv += mv
            csolved += int(ok)
        cel = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": cmv / cel, "unit": "moves/s", "cores": 8,
                               "kind": "reference", "solved": csolved,
                               "sample": f"first 4 seeds: xtax_ref {' '.join(extra)} --seed s "
                                         f"--max-seconds {args.cpu_cap}"}
    return out


def dual_roofline(kernel_ms, moves, n, greedy):
    """Algorithmic VALU work of k_anneal_dual over its measured time. Per
    proposal and column: the primal Add (multiply, add, bound check, nonzero
    change, the two exact factors: 5 int64 ops, 2 conversions and a product),
    the dual Add (shear product, add, two squares, difference: 5 fp64 ops) and
    the two column sums; with probability `greedy` the best shear's four
    products (2 conversions) and four sums. An int64 op counts 2 int32
    lane-ops, an fp64 add/mul 1 (gfx950 issues vector fp64 at the int32 rate)."""
    if not kernel_ms:
        return None
    per_col = 5 * 2 + 3 + 5 + 2 + greedy * (4 + 2 + 4)
    achieved = moves * per_col * n / (kernel_ms / 1e3) / 1e12
    peak = MI355X_INT_LANE_OPS / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
            "frac": achieved / peak, "traffic": None, "kernel": "k_anneal_dual",
            "kernel_ms": kernel_ms, "moves": moves,
            "note": "int64 + fp64 VALU issue bound; P and Q live in LDS, HBM is touched for "
                    "X^T rows and epoch boundaries"}


def bench_dual(args, eng):
    _, G = workloads.lattice(24, 24, seed=2)
    n = G.shape[0]
    p = EngineParams(replicas=4096, greedy_fraction=0.5, exchange_interval=args.exchange_interval,
                     max_seconds=args.seconds, seed=args.seed, verify=True)
    t0 = time.perf_counter()
    r = eng.run_xdual(G, None, p, DualParams())
    el = time.perf_counter() - t0
    st = r.stats
    cfg = {"workload": "dual: xdual objective on the 24x24 lattice Gram (basis U[-9,9]), 4096 "
                       "replicas, greedy 0.5, lambda 1", "n": n, "replicas": 4096}
    out = base_line(args, st["moves"] / el, "moves/s", el, cfg)
    out["dtype"] = "int64+f64"
    out.update({"solved": st["solved"], "best_score": st["best_score"],
                "start_score": st["start_score"], "c": st["c"], "epochs": st["epochs"],
                "refreshes": st["refreshes"], "dual_residual": st["dual_residual"],
                "verify_status": st["verify_status"],
                "kernel_moves_per_sec": st["moves"] / (st["kernel_ms"] / 1e3),
                "roofline": dual_roofline(st["kernel_ms"], st["moves"], n, 0.5),
                "cpu_baseline": None})
    if not args.no_cpu and os.path.exists(REF_DUAL_BIN):
        cores = os.cpu_count() or 1
        ok, mv, best = run_ref(G, ["-t", str(cores), "--seed", "1", "--max-seconds",
                                   str(args.cpu_cap)], timeout=args.cpu_cap + 30,
                               binary=REF_DUAL_BIN)
        out["cpu_baseline"] = {"value": mv / args.cpu_cap, "unit": "moves/s", "cores": cores,
                               "kind": "reference", "best_score": best, "solved": ok,
                               "sample": f"xdual_ref -t {cores} --max-seconds {args.cpu_cap}"}
    return out


def bench_copy(args, eng):
    """Calibration for the HBM traffic counters: --count device-to-device
    copies of 256 MiB (one copy kernel each), so FETCH_SIZE / WRITE_SIZE per
    dispatch can be compared with a known 256 MiB read and write."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    t0 = time.perf_counter()
    for _ in range(args.count):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # DeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    el = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    out = base_line(args, args.count * 2 * nbytes / el / 1e9, "GB/s", el,
                    {"workload": f"copy: {args.count} x 256 MiB device-to-device"})
    out["bytes_per_copy"] = nbytes
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "plain", "serve", "large", "lattice", "random", "dual",
                             "copy"])
    ap.add_argument("--count", type=int, default=None)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--t-init", type=float, default=None)
    ap.add_argument("--t-min", type=float, default=None)
    ap.add_argument("--target-fraction", type=float, default=None)
    ap.add_argument("--exchange-interval", type=int, default=2000)
    ap.add_argument("--cap-epochs", type=int, default=50)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--lookahead", type=int, default=0,
                    help="proposals of a replica evaluated at once (0: the engine's rule)")
    ap.add_argument("--cpu-seconds", type=float, default=None,
                    help="seconds per CPU baseline run (batch: 10, three runs each way)")
    ap.add_argument("--cpu-cap", type=float, default=None)
    ap.add_argument("--cpu-cap-single", type=float, default=1.0,
                    help="batch, CPU mode (ii): seconds per single-thread instance")
    ap.add_argument("--plain-count", type=int, default=64,
                    help="batch: instances also run one per launch (the plain form)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-persistent", action="store_true",
                    help="batch: captured graphs of epochs instead of one persistent launch")
    ap.add_argument("--no-tune", action="store_true",
                    help="batch: launch configurations by rule instead of the autotuner")
    ap.add_argument("--x-in-lds", action="store_true",
                    help="large: each replica's X^T in LDS for the epoch (64 lanes, 2 "
                         "columns, lookahead 4) instead of in HBM")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; instances are independent, so N GPUs "
              "would each take 1/N of the batch. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.cpu_cap is None:
        # batch: the reference gets the GPU's per-instance budget in moves
        # (replicas x cap moves = 3.2M at the defaults, about 0.05 s on 16 cores),
        # doubled; single-instance runs get a fixed 2 s
        args.cpu_cap = 0.1 if args.workload == "batch" else 2.0
    if args.cpu_seconds is None:
        args.cpu_seconds = 10.0 if args.workload == "batch" else 20.0
    ladder = {"batch": (18.0, 14.0), "plain": (18.0, 14.0), "serve": (50.0, 36.0),
              "large": (2000.0, 50.0)}.get(args.workload, (200.0, 5.0))
    if args.target_fraction is None:
        args.target_fraction = 1.0 if args.workload in ("batch", "plain", "serve") else 0.5
    if args.t_init is None:
        args.t_init = ladder[0]
    if args.t_min is None:
        args.t_min = ladder[1]
    if args.count is None:
        args.count = {"batch": 1024, "plain": 64, "serve": 3000, "random": 64, "copy": 8}.get(
            args.workload, 1)
    eng = Engine(0)
    warm(eng)
    fn = {"batch": bench_batch, "plain": bench_plain, "serve": bench_serve,
          "large": bench_large, "lattice": bench_lattice, "random": bench_random,
          "dual": bench_dual, "copy": bench_copy}[args.workload]
    out = fn(args, eng)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
