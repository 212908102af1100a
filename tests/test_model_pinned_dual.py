"""The oracle's trajectory model of the engine on xdual's objective, pinned
against the reference's own xdual runs.

The engine's dual kernel is tested bit for bit against the model
(tests/test_gpu_dual.py), so what the model decides is what the engine
decides. The model's loop body (DualModelT::step_with in
oracle/oracle_dual.cpp: pair tournaments, greedy coin, the dual best shear,
Metropolis on the combined delta, local best, stall sweep, reheat, reseed,
cooling and adaptive cooling) runs over two arithmetics and two draw sources:
the engine's double specification with its Philox words, and the reference's
long double sums with its draws (std::mt19937, libstdc++'s distributions, the
redraw loops, std::exp), its publish rule (an accepted downhill move below the
global best recomputes the score, and a diagonal global best stops the run;
congruence_anneal.hpp:309-352) and its maintenance after every move
(xdual_main.cpp:305-334). In the reference's mode it must reproduce the
reference exactly:

* G6 (tests/golden/g6_xdual_runs.npz): 19 runs of the reference's xdual binary
  (`xdual_ref -t 1`, deterministic up to the diagonal it stops at), over
  planted forms with dual weights 0 to 1 and non-default flags (sweeps,
  greedy 0 and 1, tournaments of 3, hot-row fraction 0.9, reseeds, adaptive
  cooling, cooling, forced re-inversions, residual checks at a tolerance that
  always re-inverts, --gram-rows, -X): the same "done moves=" count, best_P,
  best_Q (bit for bit) and best_X;
* G6b (tests/golden/g6b_dual_worker.npz): 48 recorded runs of the reference's
  own canneal::anneal_worker on DualObjective (one worker of a pool, tempering
  on a fixed rung or cooling, random stall, dual weight and maintenance
  settings, capped at a move count): every accepted move, sweep commit and
  reseed, the final P, Q and X^T, the global best and both scores;
* live, where the reference harness is built here: random configurations.

What differs between the two modes is only the arithmetic (the engine's
double specification, pinned against the reference's primitives by G4 in
tests/test_oracle_dual.py), the draw source (Philox, pinned by known
answers), and the engine's deliberate coordination changes (maintenance and
per-replica bests at epoch boundaries, checked there by invariants). Each
plausible mistake planted in the shared loop body changes a reference run.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# engine parameters of an xdual command line (xdual_main.cpp:463-482)
_KW = {"--sweep-fraction": ("sweep_fraction", float),
       "--target-fraction": ("target_fraction", float),
       "--greedy-fraction": ("greedy_fraction", float),
       "--stuck-threshold": ("stuck_threshold", int),
       "--moves-per-cool": ("moves_per_cool", int),
       "--target-samples": ("target_samples", int),
       "--t-init": ("t_init", float), "--t-min": ("t_min", float),
       "--reheat": ("reheat", float), "--reseed-factor": ("reseed_factor", float),
       "--cooling": ("cooling", float)}


@pytest.fixture(scope="module")
def g6():
    return np.load(os.path.join(GOLD, "g6_xdual_runs.npz"))


@pytest.fixture(scope="module")
def g6b():
    return np.load(os.path.join(GOLD, "g6b_dual_worker.npz"))


@pytest.fixture
def mutation(orc):
    yield orc.set_mutation
    orc.set_mutation(0)


def _flag(args, name, default=None):
    return args[args.index(name) + 1] if name in args else default


def _index_spec(spec):
    out = []
    for part in spec.split(","):
        if ".." in part:
            lo, hi = part.split("..")
            out += list(range(int(lo), int(hi) + 1))
        else:
            out.append(int(part))
    return out


def _case(g6, c):
    """(name, P0, X0, seed, run keywords, expected) of G6 run c: the working
    Gram starts at X^T A X with -X (xdual_main.cpp:536-538), else at A."""
    args = str(g6[f"c{c}_args"]).split()
    A0 = g6[f"c{c}_A0"]
    n = A0.shape[0]
    X0 = g6[f"c{c}_X0"] if g6[f"c{c}_X0"].size else np.eye(n, dtype=np.int64)
    P0 = (X0.T.astype(object) @ A0.astype(object) @ X0.astype(object)).astype(np.int64)
    kw = {_KW[args[k]][0]: _KW[args[k]][1](args[k + 1])
          for k in range(len(args) - 1) if args[k] in _KW}
    if "--adaptive-cooling" in args:
        kw["adaptive_cooling"] = 1
    kw.update(lam=float(_flag(args, "--lambda", 1.0)),
              dual_refresh=int(_flag(args, "--dual-refresh", 200000)),
              dual_check=int(_flag(args, "--dual-check", 25000)),
              dual_tol=float(_flag(args, "--dual-tol", 1e-6)))
    if "--gram-rows" in args:
        kw["active"] = _index_spec(_flag(args, "--gram-rows"))
    want = dict(moves=int(g6[f"c{c}_moves"]), P=g6[f"c{c}_P"], Q=g6[f"c{c}_Q"],
                X=g6[f"c{c}_X"])
    return str(g6[f"c{c}_name"]), P0, X0, int(_flag(args, "--seed")), kw, want


def _run(orc, P0, X0, seed, kw, **extra):
    return orc.dual_ref_run(P0, X0, orc.ref_worker_seed(seed, 0), **kw, **extra)


def _same(m, want):
    return (m["solved"] == 1 and m["moves"] == want["moves"]
            and np.array_equal(m["P"], want["P"]) and np.array_equal(m["Q"], want["Q"])
            and np.array_equal(m["Xt"].T, want["X"]))


def test_reproduces_every_reference_xdual_run(orc, g6):
    runs = int(g6["runs"])
    assert runs == 19
    refreshes = checks = sweeps = reseeds = 0
    names = set()
    for c in range(runs):
        name, P0, X0, seed, kw, want = _case(g6, c)
        m = _run(orc, P0, X0, seed, kw, log_cap=1 << 16)
        assert _same(m, want), (name, m["solved"], m["moves"], want["moves"])
        # the best is the input's congruence by the returned transform (with -X
        # the transform includes the start's, xdual_main.cpp:533-535)
        X, A0 = m["Xt"].T.astype(object), g6[f"c{c}_A0"].astype(object)
        assert np.array_equal((X.T @ A0 @ X).astype(np.int64), m["P"]), name
        refreshes += m["refreshes"]
        checks += m["checks"]
        sweeps += int((m["log"]["flags"] == 1).sum())
        reseeds += int((m["log"]["flags"] == 2).sum())
        names.add(name)
    # the fixtures exercise re-inversion, residual checks, sweeps and reseeds
    assert refreshes > 50 and checks > 50 and sweeps > 0 and reseeds > 50
    assert {"gram_rows", "p6_start_X", "p7_check", "p7_refresh", "p7_lambda0"} <= names


def test_the_reference_run_reinverts_on_a_residual_above_tolerance(orc, g6):
    """At a tolerance of 1e-15 nearly every residual check re-inverts
    (xdual_main.cpp:316-334); with a refresh period, every period re-inverts
    besides."""
    cases = {_case(g6, c)[0]: _case(g6, c) for c in range(int(g6["runs"]))}
    _, P0, X0, seed, kw, _ = cases["p7_check"]
    m = _run(orc, P0, X0, seed, kw)
    assert m["checks"] == m["moves"] // 5000 and m["refreshes"] >= 0.9 * m["checks"]
    _, P0, X0, seed, kw, want = cases["p7_refresh"]
    m = _run(orc, P0, X0, seed, kw)
    assert m["refreshes"] >= want["moves"] // 20000


@pytest.mark.parametrize("which", [1, 2, 5, 6])
def test_planted_mistakes_change_a_reference_run(orc, g6, mutation, which):
    """1 strict auto-accept, 2 reversed hot-row tournament, 5 Metropolis
    exponent sign, 6 the dual update committed to row j instead of i: each
    changes at least one reference run (runs capped at twice their length)."""
    mutation(which)
    order = sorted(range(int(g6["runs"])), key=lambda c: int(g6[f"c{c}_moves"]))
    for c in order:
        name, P0, X0, seed, kw, want = _case(g6, c)
        if want["moves"] > 700_000:
            continue
        m = _run(orc, P0, X0, seed, kw, max_moves=2 * want["moves"] + 1000)
        if not _same(m, want):
            return
    pytest.fail(f"mutation {which} reproduces every reference run")


# -------------------------------------------------- the reference's worker loop
def _same_trace(orc, P0, X0, act, dp, ip, want):
    from oracle import refharness
    m = refharness.dual_model(P0, X0, dp, ip, act)
    got = np.stack([m["log"]["move"], m["log"]["i"], m["log"]["j"], m["log"]["s"],
                    m["log"]["flags"]], 1).astype(np.int64)
    # the recorder reports its global best's diagonality; the model, whether
    # the run stopped at a diagonal best (the same unless the start is diagonal)
    gb_diag = int(orc.offdiag_nonzero(m["P"]) == 0)
    return (m["moves"] == want["sc"][0] and gb_diag == want["sc"][1]
            and m["score"] == want["dsc"][0] and m["gb_score"] == want["dsc"][1]
            and m["c"] == want["dsc"][2] and got.shape == want["log"].shape
            and np.array_equal(got, want["log"])
            and all(np.array_equal(m[k], want[k]) for k in ("Pf", "Qf", "Xtf", "P", "Q", "Xt")))


def test_reproduces_reference_dual_worker_traces(orc, g6b):
    T = int(g6b["traces"])
    assert T == 48
    temper = cool = 0
    for t in range(T):
        want = {k: g6b[f"t{t}_{k}"] for k in ("Pf", "Qf", "Xtf", "P", "Q", "Xt", "log", "sc",
                                               "dsc")}
        ip = g6b[f"t{t}_ip"]
        assert _same_trace(orc, g6b[f"t{t}_P0"], g6b[f"t{t}_X0"], g6b[f"t{t}_act"],
                           g6b[f"t{t}_dp"], ip, want), t
        if ip[3] and ip[5] >= 2:
            temper += 1
        else:
            cool += 1
    assert temper >= 10 and cool >= 10
    flags = np.concatenate([g6b[f"t{t}_log"][:, 4] for t in range(T)])
    assert (flags == 1).sum() > 50 and (flags == 2).sum() > 50


def test_live_against_reference_dual_worker(orc):
    from oracle import refharness
    if not refharness.dual_available():
        pytest.skip("the reference harness is built only where the reference sources are")
    import sys
    sys.path.insert(0, GOLD)
    from gen_golden import g6b_configs
    rng = np.random.default_rng(78)
    for t, (P0, X0, act, dp, ip) in enumerate(g6b_configs(rng, 60)):
        r = refharness.dual_trace(P0, X0, dp, ip, act)
        want = dict(r, log=np.stack([r["log"]["move"], r["log"]["i"], r["log"]["j"],
                                     r["log"]["s"], r["log"]["flags"]], 1).astype(np.int64),
                    sc=np.array([r["moves"], r["solved"]]),
                    dsc=np.array([r["score"], r["gb_score"], r["c"]]))
        assert _same_trace(orc, P0, X0, act, dp, ip, want), t


@pytest.mark.parametrize("which", [2, 5])
def test_planted_mistakes_change_a_worker_trace(orc, g6b, mutation, which):
    """In both temperature modes (a reversed hot-row tournament, a wrong
    Metropolis sign)."""
    mutation(which)
    changed = set()
    for t in range(int(g6b["traces"])):
        ip = g6b[f"t{t}_ip"]
        mode = "temper" if ip[3] and ip[5] >= 2 else "cool"
        if mode in changed:
            continue
        want = {k: g6b[f"t{t}_{k}"] for k in ("Pf", "Qf", "Xtf", "P", "Q", "Xt", "log", "sc",
                                               "dsc")}
        if not _same_trace(orc, g6b[f"t{t}_P0"], g6b[f"t{t}_X0"], g6b[f"t{t}_act"],
                           g6b[f"t{t}_dp"], ip, want):
            changed.add(mode)
    assert changed == {"temper", "cool"}
