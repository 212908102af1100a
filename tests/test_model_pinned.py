"""The oracle's trajectory model pinned against the reference's own annealer.

The engine is tested bit for bit against the model (tests/test_gpu_*.py), so
what the model decides is what the engine decides. The model's loop body
(Model::step_with in oracle/oracle.cpp: pair tournaments, greedy coin, shear
choice, Metropolis, local best, stall sweep, reheat, reseed, cooling and
adaptive cooling) is shared by two draw sources: the engine's Philox words and
the reference's draws (std::mt19937, libstdc++'s distributions, the redraw
loops, std::exp). Run on the reference's draws with the reference's publish
rule, it must reproduce the reference exactly:

* G2 and G5 runs of the reference binary (`xtax_ref -t 1`, deterministic up
  to the diagonal it stops at): the same "done moves=" count, best_A and
  best_X, over the sample matrix, planted forms and non-default flags
  (sweeps, adaptive cooling, tournament size, greedy 0 and 1, reseeds,
  cooling, uniform pairs, --gram-rows, -X);
* G5 traces of canneal::anneal_worker itself (one worker of a pool, recorded
  through an objective wrapper): every accepted move, sweep commit and reseed,
  the final state and the global best, in both temperature modes; G5b the
  same with tournaments of 17 to 257 samples;
* live, where the reference harness is built here: random configurations.

What the reference decides with wall-clock timing (the 5 ms exchange monitor,
congruence_anneal.hpp:502-555) is pinned by closed forms and statistics: the
ladder (:454-468) and the exchange rule (:520-551) and Metropolis test (:236)
against min(1, e^x) per bucket. Each check is shown to fail when a plausible
mistake is planted in the model (orc_set_mutation).
"""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def g2():
    return np.load(os.path.join(GOLD, "g2_end_to_end.npz"))


@pytest.fixture(scope="module")
def g5():
    return np.load(os.path.join(GOLD, "g5_worker.npz"))


@pytest.fixture(scope="module")
def g5b():
    return np.load(os.path.join(GOLD, "g5b_samples.npz"))


@pytest.fixture
def mutation(orc):
    yield orc.set_mutation
    orc.set_mutation(0)


def _flag(args, name, default=None):
    return args[args.index(name) + 1] if name in args else default


def _g2_runs(g2):
    """(A0, seed, moves, best_A, best_X) of every deterministic G2 run."""
    for t in range(len(g2["names"])):
        args = str(g2["args"][t]).split()
        if _flag(args, "-t") != "1" or not g2["meta"][t][1]:
            continue
        yield (g2[f"r{t}_A0"], int(_flag(args, "--seed")), int(g2["meta"][t][2]),
               g2[f"r{t}_A"], g2[f"r{t}_X"])


def _cli_kw(args):
    """Engine parameters of a reference command line (main.cpp:578-624)."""
    kw = {}
    for flag, key, typ in [("--sweep-fraction", "sweep_fraction", float),
                           ("--stuck-threshold", "stuck_threshold", int),
                           ("--moves-per-cool", "moves_per_cool", int),
                           ("--target-samples", "target_samples", int),
                           ("--target-fraction", "target_fraction", float),
                           ("--greedy-fraction", "greedy_fraction", float),
                           ("--t-init", "t_init", float), ("--t-min", "t_min", float),
                           ("--reheat", "reheat", float), ("--reseed-factor", "reseed_factor", float),
                           ("--cooling", "cooling", float)]:
        if flag in args:
            kw[key] = typ(_flag(args, flag))
    if "--adaptive-cooling" in args:
        kw["adaptive_cooling"] = 1
    return kw


def _index_spec(spec):
    out = []
    for part in spec.split(","):
        if ".." in part:
            lo, hi = part.split("..")
            out += list(range(int(lo), int(hi) + 1))
        else:
            out.append(int(part))
    return out


def _reference_run(orc, A0, seed, kw=None, active=None, X0=None):
    """run_annealer with one thread through the model's reference-draw mode."""
    n = A0.shape[0]
    X0 = np.eye(n, dtype=np.int64) if X0 is None else X0
    A_start = (X0.T.astype(object) @ A0.astype(object) @ X0.astype(object)).astype(np.int64)
    return orc.ref_worker_run(A_start, X0, orc.ref_worker_seed(seed, 0), threads=1,
                              active=active, **(kw or {}))


# ------------------------------------------------------ the reference binary
def test_reproduces_every_deterministic_reference_run(orc, g2):
    runs = list(_g2_runs(g2))
    assert len(runs) == 13
    for A0, seed, moves, bA, bX in runs:
        r = _reference_run(orc, A0, seed)
        assert r["solved"] == 1 and r["moves"] == moves, (seed, r["moves"], moves)
        assert np.array_equal(r["gbA"], bA) and np.array_equal(r["gbXt"].T, bX)


def test_reproduces_reference_runs_with_flags(orc, g5):
    k = int(g5["cli_runs"])
    assert k >= 10
    for c in range(k):
        args = str(g5[f"c{c}_args"]).split()
        A0 = g5[f"c{c}_A0"]
        X0 = g5[f"c{c}_X0"] if g5[f"c{c}_X0"].size else None
        act = _index_spec(_flag(args, "--gram-rows")) if "--gram-rows" in args else None
        r = _reference_run(orc, A0, int(_flag(args, "--seed")), _cli_kw(args), act, X0)
        name = str(g5[f"c{c}_name"])
        assert r["solved"] == 1 and r["moves"] == int(g5[f"c{c}_moves"]), (name, r["moves"])
        assert np.array_equal(r["gbA"], g5[f"c{c}_A"]), name
        assert np.array_equal(r["gbXt"].T, g5[f"c{c}_X"]), name


# -------------------------------------------------- the reference's worker loop
def _trace_case(g5, t):
    lg = g5[f"t{t}_log"]
    return (g5[f"t{t}_A"], g5[f"t{t}_X0"], g5[f"t{t}_act"], g5[f"t{t}_dp"], g5[f"t{t}_ip"],
            dict(A=g5[f"t{t}_Af"], Xt=g5[f"t{t}_Xtf"], gbA=g5[f"t{t}_gbA"], gbXt=g5[f"t{t}_gbXt"],
                 log=lg, sc=g5[f"t{t}_sc"]))


def _same_trace(orc, A, X0, act, dp, ip, want):
    from oracle import refharness
    m = refharness.model(A, X0, dp, ip, act)
    got = np.stack([m["log"]["move"], m["log"]["i"], m["log"]["j"], m["log"]["s"],
                    m["log"]["flags"]], 1).astype(np.int64)
    return (m["moves"] == want["sc"][0] and m["score"] == want["sc"][1]
            and m["gb_score"] == want["sc"][2] and got.shape == want["log"].shape
            and np.array_equal(got, want["log"]) and np.array_equal(m["A"], want["A"])
            and np.array_equal(m["Xt"], want["Xt"]) and np.array_equal(m["gbA"], want["gbA"])
            and np.array_equal(m["gbXt"], want["gbXt"]))


def test_reproduces_reference_worker_traces(orc, g5):
    T = int(g5["traces"])
    temper = cool = 0
    for t in range(T):
        A, X0, act, dp, ip, want = _trace_case(g5, t)
        assert _same_trace(orc, A, X0, act, dp, ip, want), t
        if ip[3] and ip[5] >= 2:
            temper += 1
        else:
            cool += 1
        assert want["sc"][0] == ip[8] or want["sc"][3] == 1  # ran to the cap or solved
    assert temper >= 10 and cool >= 10
    # the fixtures exercise sweeps, reseeds and both outcomes
    flags = np.concatenate([g5[f"t{t}_log"][:, 4] for t in range(T)])
    assert (flags == 1).any() and (flags == 2).any()


def test_reproduces_reference_worker_traces_large_tournaments(orc, g5b):
    """Tournaments of more than 16 samples (the engine's wide mode draws them)."""
    T = int(g5b["traces"])
    sizes = set()
    for t in range(T):
        A, X0, act, dp, ip, want = _trace_case(g5b, t)
        assert ip[2] > 16
        sizes.add(int(ip[2]))
        assert _same_trace(orc, A, X0, act, dp, ip, want), t
    assert len(sizes) >= 6


def test_slots_past_16_leave_the_first_16_alone(orc):
    """The engine's draw counters: slot c of move m at word 64 (c / 16) +
    16 ((m - 1) mod 4) + c mod 16, so the slots every tournament of at most 16
    samples uses keep their counters, and no two (move, slot) pairs of a
    replica share one (brute force over four moves and 1024 slots)."""
    words = {}
    for t in range(4):
        for c in range(1024):
            w = 64 * (c >> 4) + 16 * t + (c & 15)
            if c < 16:
                assert w == 16 * t + c
            assert w not in words, (t, c, words.get(w))
            words[w] = (t, c)


def test_live_against_reference_worker(orc):
    from oracle import refharness
    if not refharness.available():
        pytest.skip("the reference harness is built only where the reference sources are")
    import sys
    sys.path.insert(0, GOLD)
    from gen_golden import g5_configs
    rng = np.random.default_rng(77)
    for t, (A, X0, act, dp, ip) in enumerate(g5_configs(rng, 150)):
        r = refharness.trace(A, X0, dp, ip, act)
        want = dict(A=r["A"], Xt=r["Xt"], gbA=r["gbA"], gbXt=r["gbXt"],
                    log=np.stack([r["log"]["move"], r["log"]["i"], r["log"]["j"],
                                  r["log"]["s"], r["log"]["flags"]], 1).astype(np.int64),
                    sc=np.array([r["moves"], r["score"], r["gb_score"], r["solved"]]))
        assert _same_trace(orc, A, X0, act, dp, ip, want), t


# ---------------------------------------------------- planted mistakes are seen
@pytest.mark.parametrize("which", [1, 2, 5])
def test_planted_step_mistakes_break_the_reference_runs(orc, g2, mutation, which):
    """Strict auto-accept, a reversed hot-row tournament, a wrong Metropolis
    sign: each changes a reference run's outcome."""
    mutation(which)
    diverged = 0
    for A0, seed, moves, bA, bX in list(_g2_runs(g2))[:12]:
        r = _reference_run(orc, A0, seed, dict(max_moves=4 * moves + 1000))
        if not (r["moves"] == moves and np.array_equal(r["gbA"], bA)):
            diverged += 1
    assert diverged >= 6


def test_planted_diversity_mistake_breaks_the_worker_traces(orc, g5, mutation):
    mutation(3)
    diverged = total = 0
    for t in range(int(g5["traces"])):
        A, X0, act, dp, ip, want = _trace_case(g5, t)
        if ip[5] < 2 or dp[8] == 0 or (dp[5] in (0.0, 1.0) and dp[6] in (0.0, 1.0)):
            continue  # no diversity offset in play
        total += 1
        diverged += not _same_trace(orc, A, X0, act, dp, ip, want)
    assert total >= 5 and diverged >= total // 2


# -------------------------------------------------------------- closed forms
@pytest.mark.parametrize("top,t_min,W", [(50.0, 0.5, 8), (18.0, 14.0, 32), (2000.0, 50.0, 4096),
                                          (3.0, 1e-3, 256), (0.2, 5.0, 16), (7.0, 7.0, 5)])
def test_ladder_closed_form(orc, top, t_min, W):
    """congruence_anneal.hpp:454-468: geometric from min(t_min, top) to top."""
    t = np.zeros(W)
    from oracle.pyoracle import lib
    import ctypes as C
    lib().orc_ladder(top, t_min, W, t.ctypes.data_as(C.POINTER(C.c_double)))
    bot = min(t_min, top)
    assert t[0] == bot
    ratio = t[1] / t[0]
    assert np.all(np.abs(t[1:] / t[:-1] - ratio) <= 4e-16 * ratio)
    assert abs(t[-1] - top) <= W * 2.3e-16 * top
    assert np.all(np.diff(t) >= 0)
    t1 = np.zeros(1)
    lib().orc_ladder(top, t_min, 1, t1.ctypes.data_as(C.POINTER(C.c_double)))
    assert t1[0] == top


def _binomial_ok(k, N, p, z=5.0):
    return abs(k - N * p) <= z * math.sqrt(max(N * p * (1 - p), 1e-12)) + 1


def test_metropolis_acceptance_statistics(orc):
    """The model's Metropolis test accepts an uphill move of change d at
    temperature T with probability e^(-d/T), every move with d <= 0."""
    rng = np.random.default_rng(11)
    N = 4000
    words = rng.integers(0, 1 << 32, (N, 2), dtype=np.uint64)
    for d, T in [(-5, 1.0), (0, 0.5), (1, 1.0), (3, 2.0), (10, 4.0), (40, 10.0), (7, 100.0),
                 (200, 50.0)]:
        k = sum(orc.metropolis(d, T, int(w[0]), int(w[1])) for w in words)
        p = 1.0 if d <= 0 else math.exp(-d / T)
        if d <= 0:
            assert k == N
        else:
            assert _binomial_ok(k, N, p), (d, T, k / N, p)


def test_exchange_acceptance_statistics(orc):
    """Adjacent rungs swap with probability min(1, e^arg), arg = (1/T_r -
    1/T_{r+1})(S_r - S_{r+1}), judged for the occupant carried up a rung: over
    many epochs of a two-rung ladder, per bucket of arg."""
    temps = np.array([1.0, 2.0])
    rng = np.random.default_rng(5)
    for dS in [-8.0, -3.0, -1.0, -0.25, 0.0, 0.5, 4.0]:
        # rung 0 holds a configuration dS worse (higher) than rung 1's
        scores = np.array([100.0 + dS, 100.0])
        arg = (1 / temps[0] - 1 / temps[1]) * (scores[0] - scores[1])
        N = 3000
        acc = 0
        for e in range(N):
            _, a = orc.exchange_sweep(temps, scores, [0, 1], int(rng.integers(1 << 62)), e)
            acc += a
        p = min(1.0, math.exp(arg))
        if arg >= 0:
            assert acc == N, (dS, acc)
        else:
            assert _binomial_ok(acc, N, p), (dS, acc / N, p)


def test_exchange_carries_a_configuration_up_the_ladder(orc):
    """The sweep runs bottom up and a swapped configuration is compared again
    at the next pair (congruence_anneal.hpp:533-549): a far worse
    configuration on rung 0 of a flat-score ladder climbs to the top in one
    sweep."""
    R = 6
    temps = np.geomspace(1.0, 32.0, R)
    scores = np.array([1000.0] + [0.0] * (R - 1))
    rungs, acc = orc.exchange_sweep(temps, scores, np.arange(R), 3, 0)
    assert rungs[0] == R - 1 and acc >= R - 1


def test_planted_exchange_mistake_fails_the_statistics(orc, mutation):
    mutation(4)
    temps = np.array([1.0, 2.0])
    scores = np.array([97.0, 100.0])  # arg = -1.5: accepted 22% of the time
    acc = sum(orc.exchange_sweep(temps, scores, [0, 1], 9, e)[1] for e in range(2000))
    assert not _binomial_ok(acc, 2000, math.exp(-1.5))


def test_planted_metropolis_mistake_fails_the_statistics(orc, mutation):
    # (a strict auto-accept, d < 0, is invisible here -- u < e^0 always holds --
    # and is caught by the reference runs, whose stream it shifts by a draw)
    rng = np.random.default_rng(3)
    words = rng.integers(0, 1 << 32, (3000, 2), dtype=np.uint64)
    mutation(5)
    k = sum(orc.metropolis(3, 2.0, int(w[0]), int(w[1])) for w in words)
    assert not _binomial_ok(k, 3000, math.exp(-1.5))
