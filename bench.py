"""Benchmark of the gfx950 xtax annealer (one GPU).

Workload (default, `--workload batch`): 1024 planted instances, n ~ U{8..48},
A = U^T D U with D = diag(+-U{1..100}) and U a product of 3n random +-1 shears
(xtax_amd.workloads.batch_config), 32 tempering replicas each on a geometric
ladder T in [5, 200], greedy fraction 0.5, exchange every 2000 moves, all
instances resident in one captured launch sequence; an instance retires when a
replica reaches a diagonal or after `--cap-epochs` epochs (its cap). One step =
one pass over the whole batch. Every solved result is re-verified exactly on
the host (X^T A0 X = A, A diagonal, |det X| = 1) and only verified ones count.

Prints one JSON line: metric/value/unit (exact-diagonalized matrices per
second), ms_per_step, solve latency percentiles, shear moves per second, the
roofline of the dominant kernel (k_anneal) and a CPU baseline: the reference's
own xtax binary (oracle/_ref/xtax_ref, compiled from the reference sources) run
on a bounded sample of the same instances on this host's cores.

Other workloads: `--workload plain` (the same instances one per launch),
`serve` (Poisson request stream, continuous batching), `large` (n = 80
planted, 4096 replicas), `lattice` (24x24 Gram, 256 replicas),
`random` (n = 10 random symmetric, 8 replicas).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from xtax_amd import Engine, EngineParams, workloads  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "xtax_ref")
MI355X_INT_LANE_OPS = 256 * 64 * 2.4e9  # CUs * lanes/clk/CU * clock: 39.3 T int32 lane-ops/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, float), q)) if len(xs) else None


def batch_params(args):
    return EngineParams(replicas=args.replicas, t_init=args.t_init, t_min=args.t_min,
                        greedy_fraction=0.5, exchange_interval=args.exchange_interval,
                        max_epochs=args.cap_epochs, seed=args.seed, verify=True)


def run_batch(eng, mats, args, seeds):
    p = batch_params(args)
    res, bst = eng.batch_solve(mats, p, seeds=seeds)
    return res, bst


def cpu_baseline(mats, args, budget_s):
    """The reference binary on the first instances, one at a time with every
    core (-t nproc, same ladder), each capped at --cpu-cap seconds, until the
    budget is spent."""
    if not os.path.exists(REF_BIN):
        return None
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    solved = moves = done = 0
    lat = []
    with tempfile.TemporaryDirectory() as td:
        for q, A in enumerate(mats):
            if time.perf_counter() - t0 > budget_s:
                break
            path = os.path.join(td, "A.csv")
            with open(path, "w") as f:
                for row in A:
                    f.write(",".join(str(int(v)) for v in row) + "\n")
            t1 = time.perf_counter()
            cp = subprocess.run([REF_BIN, "-A", path, "-t", str(cores), "--seed", str(q + 1),
                                 "--t-init", str(args.t_init), "--t-min", str(args.t_min),
                                 "--greedy-fraction", "0.5", "--max-seconds", str(args.cpu_cap)],
                                cwd=td, capture_output=True, text=True)
            dt = time.perf_counter() - t1
            done += 1
            for line in cp.stdout.splitlines():
                if line.startswith("done moves="):
                    moves += int(line.split()[1].split("=")[1])
            if "matrix is diagonal" in cp.stdout:
                solved += 1
                lat.append(dt)
    el = time.perf_counter() - t0
    return {"value": solved / el, "unit": "matrices/s", "cores": cores, "kind": "reference",
            "sample": f"first {done} of the {len(mats)} batch instances, one at a time, "
                      f"xtax_ref -t {cores} --t-init {args.t_init} --t-min {args.t_min} "
                      f"--max-seconds {args.cpu_cap}; {solved} solved in {el:.1f} s",
            "moves_per_sec": moves / el, "p50_latency_s": pct(lat, 50)}


def roofline(bst, moves, mean_n, args):
    """Algorithmic int64 lane-ops of the annealer kernel over its measured time.

    Per proposal the method evaluates 1 Add (6n int64 ops: mul, add, bound,
    2x abs, accumulate, over the n columns) plus, with probability
    greedy_fraction, up to 5 candidate shears (5 * 6n) and the median search
    (9 weighted counts, 3n each). An int64 op is 2 int32 lane-ops."""
    ops_per_move = 6 * mean_n * (1 + 0.5 * 5) + 0.5 * 9 * 3 * mean_n
    lane_ops = moves * ops_per_move * 2
    t = bst["kernel_ms"] / 1e3
    if t <= 0:
        return None
    achieved = lane_ops / t / 1e12
    peak = MI355X_INT_LANE_OPS / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
            "frac": achieved / peak, "traffic": None,
            "note": "integer VALU issue bound (no dense contraction, HBM touched only at "
                    "epoch boundaries); algorithmic int32 lane-ops per k_anneal second"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "plain", "serve", "large", "lattice", "random"])
    ap.add_argument("--count", type=int, default=1024)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--t-init", type=float, default=200.0)
    ap.add_argument("--t-min", type=float, default=5.0)
    ap.add_argument("--exchange-interval", type=int, default=2000)
    ap.add_argument("--cap-epochs", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-cap", type=float, default=2.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; instances are independent, so N GPUs "
              "would each take 1/N of the batch (see DESIGN.md). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    log(f"generating workload {args.workload} ...")
    t = time.perf_counter()
    mats = workloads.batch_config(count=args.count, seed=7)
    seeds = [args.seed * 1_000_003 + q for q in range(len(mats))]
    log(f"  {len(mats)} instances in {time.perf_counter() - t:.1f}s")
    mean_n = float(np.mean([a.shape[0] for a in mats]))

    eng = Engine(0)
    # device/context warm-up (not a step): one tiny solve
    tri = 2 * np.eye(4, dtype=np.int64) + np.eye(4, k=1, dtype=np.int64) + np.eye(4, k=-1, dtype=np.int64)
    eng.run_annealer(tri, None, EngineParams(replicas=2, max_epochs=1))
    for w in range(args.warmup):
        log(f"warmup step {w}")
        run_batch(eng, mats, args, seeds)

    solved = verified = 0
    moves = 0
    lat = []
    kernel_ms = 0.0
    t0 = time.perf_counter()
    last = None
    for k in range(args.steps):
        res, bst = run_batch(eng, mats, args, seeds)
        last = bst
        for A, X, d in res:
            if d["solved"]:
                solved += 1
                if d["verify_status"] == 0:
                    verified += 1
                    lat.append(d["retire_s"] - d["submit_s"])
        moves += bst["moves"]
        kernel_ms += bst["kernel_ms"]
        log(f"step {k}: solved {bst['solved']}/{bst['instances']} verified {bst['verified']} "
            f"epochs {bst['epochs']} {bst['seconds']:.2f}s kernel {bst['kernel_ms']:.0f}ms "
            f"moves/s {bst['moves'] / bst['seconds']:.3e}")
    el = time.perf_counter() - t0
    eng.close()

    value = verified / el
    out = {
        "metric": "exact-diagonalized matrices/sec + p50/p99 solve latency; shear moves/sec",
        "value": value,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"batch: {args.count} planted instances n~U{{8..48}}, "
                               f"D=+-U{{1..100}}, 3n shears, {args.replicas} replicas, "
                               f"T in [{args.t_min},{args.t_init}], greedy 0.5, exchange "
                               f"{args.exchange_interval}, cap {args.cap_epochs} epochs",
                   "instances": args.count, "replicas": args.replicas,
                   "cap_moves_per_replica": args.cap_epochs * args.exchange_interval},
        "solved": solved,
        "verified": verified,
        "solved_fraction": solved / (args.count * args.steps),
        "p50_latency_s": pct(lat, 50),
        "p95_latency_s": pct(lat, 95),
        "p99_latency_s": pct(lat, 99),
        "moves_per_sec": moves / el,
        "kernel_moves_per_sec": moves / (kernel_ms / 1e3) if kernel_ms else None,
        "roofline": roofline({"kernel_ms": kernel_ms}, moves, mean_n, args),
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(mats, args, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
