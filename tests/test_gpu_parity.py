"""Parity of the gfx950 engine (through the C ABI) with the oracle.

1. Building blocks: Philox words and exp bits; try_add / score delta / nonzero
   delta / best shear / transform bound on every golden matrix and on random
   (i, j, s) queries, for every lane-group layout (n <= 16, 32, 64, 128).
2. Trajectories: a plain solve for a fixed number of epochs ends in exactly the
   state the oracle's trajectory model reaches (every replica's A, X^T, score,
   nonzero count, move count, best, rung, accepts) with identical accepted-move
   logs; each log replays through the restated reference primitives to that
   replica's final state.
3. Results: every reported A, X passes the exact verifier; the batched and
   served forms return the same result for an instance as the plain form
   (alone vs in the mix).
4. Launch configurations: every (lanes, columns, LDS layout, waves per
   workgroup) of an order returns the rule's result; the autotuner keeps its
   fastest measurement.
"""
import os
import sys

import numpy as np
import pytest

from xtax_amd import EngineParams, workloads

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------- building blocks
def test_rng_and_exp_bits(eng, orc):
    rng = np.random.default_rng(1)
    ck = rng.integers(0, 1 << 32, (257, 6), dtype=np.uint64).astype(np.uint32)
    ck[0] = 0
    xs = np.concatenate([-rng.exponential(5.0, 2000), [0.0, -1e-300, -39.9999, -40.0, -41.0, -1e9]])
    words, ex = eng.probe_rng(ck, xs)
    for t in range(len(ck)):
        want = orc.philox(ck[t, :4].tolist(), ck[t, 4:].tolist())
        assert words[t].tolist() == want, t
    for t, x in enumerate(xs):
        assert ex[t] == orc.xexp(x), (x, ex[t], orc.xexp(x))  # bit-exact


def _check_probe(eng, orc, mats, xts, queries):
    out = eng.probe_moves(np.array(mats), np.array(xts), queries)
    for q, (mi, i, j, s) in enumerate(queries):
        A, Xt = mats[mi], xts[mi]
        ok, d, dnz, _, _ = orc.try_add(A, Xt, i, j, s)
        okt, _ = orc.add_transform_row(Xt, i, j, s)
        got = out[q].tolist()
        assert got[0] == int(ok), (mi, i, j, s)
        if ok:
            assert got[1] == d and got[2] == dnz, (mi, i, j, s, got, d, dnz)
        b = orc.best_shear_l1(A, i, j)
        assert got[3] == b, (mi, i, j, got[3])
        assert got[4] == int(okt)
        if b != 0:  # the annealer's evaluation of the greedy shear (known off sum)
            okb, db, _, _, _ = orc.try_add(A, Xt, i, j, b)
            if okb:
                assert got[5] == db, (mi, i, j, b, got[5], db)


def test_move_primitives_on_golden_matrices(eng, orc):
    g = np.load(os.path.join(GOLD, "g1_primitives.npz"))
    ns = sorted({int(k.split("_")[0][1:]) for k in g.files if k.startswith("n")})
    for n in ns:
        mats, xts = list(g[f"n{n}_A"]), list(g[f"n{n}_Xt"])
        queries = [tuple(c) for c in g[f"n{n}_cases"].tolist()]
        # add every ordered pair of the first matrices with their best-shear s
        for mi in range(min(2, len(mats))):
            for i in range(n):
                for j in range(n):
                    if i != j:
                        queries.append((mi, i, j, int(g[f"n{n}_best_all"][mi][i][j]) or 1))
        say(f"probe golden n={n} queries={len(queries)}")
        _check_probe(eng, orc, mats, xts, queries)


@pytest.mark.parametrize("n", [2, 3, 7, 16, 17, 31, 32, 33, 47, 64, 65, 80, 96])
def test_move_primitives_random(eng, orc, n):
    rng = np.random.default_rng(n)
    mats, xts = [], []
    for kind in range(6):
        scale = [3, 40, 3000, 1 << 30, 1 << 46, 1 << 47][kind]
        A = rng.integers(-scale, scale + 1, (n, n)) * (rng.random((n, n)) < 0.85)
        A = np.triu(A) + np.triu(A, 1).T
        mats.append(A.astype(np.int64))
        X = rng.integers(-scale, scale + 1, (n, n)).astype(np.int64)
        xts.append(X)
    queries = []
    for _ in range(400):
        mi = int(rng.integers(0, len(mats)))
        i = int(rng.integers(0, n))
        j = int(rng.integers(0, n - 1))
        j += j >= i
        s = int(rng.choice([-64, -2, -1, 1, 2, 64, int(rng.integers(-64, 65))]))
        queries.append((mi, i, j, s))
    say(f"probe random n={n}")
    _check_probe(eng, orc, mats, xts, queries)


# ----------------------------------------------------------------- trajectories
def _both(**kw):
    """EngineParams and oracle params from one set of values."""
    from oracle import pyoracle
    ep = EngineParams(**kw)
    op = pyoracle.make_params(**{k: v for k, v in kw.items()
                                 if k in pyoracle.DEFAULTS})
    return ep, op


CASES = [
    # name, matrix, params
    ("10x10-temper", "ten", dict(replicas=8, seed=3, exchange_interval=300, t_init=50, t_min=0.5)),
    ("10x10-cool", "ten", dict(replicas=4, seed=5, tempering=False, exchange_interval=200,
                               stuck_threshold=150, sweep_fraction=0.7, adaptive_cooling=True,
                               moves_per_cool=50)),
    ("rand20-temper", "rand20", dict(replicas=6, seed=9, exchange_interval=250, t_init=40,
                                     t_min=0.5, target_samples=3, stuck_threshold=100,
                                     sweep_fraction=0.5)),
    ("lattice24", "lat24", dict(replicas=5, seed=11, exchange_interval=200, greedy_fraction=0.5,
                                sweep_fraction=0.05, stuck_threshold=120, target_samples=16)),
    ("planted40", "p40", dict(replicas=4, seed=13, exchange_interval=150, t_init=200, t_min=5,
                              greedy_fraction=1.0, worker_diversity=0.0)),
    ("planted72", "p72", dict(replicas=3, seed=17, exchange_interval=120, t_init=300, t_min=5,
                              greedy_fraction=0.7)),
    ("rand12-cool-reseed", "rand12", dict(replicas=3, seed=19, tempering=False,
                                           exchange_interval=100, stuck_threshold=60,
                                           reseed_factor=1.0, t_init=30)),
]


def _matrix(name):
    if name == "ten":
        return np.load(os.path.join(GOLD, "g2_end_to_end.npz"))["r0_A0"]
    if name == "rand20":
        return workloads.random_symmetric(20, -20, 20, seed=4)
    if name == "rand12":
        return workloads.random_symmetric(12, -20, 20, seed=6)
    if name == "lat24":
        return workloads.lattice(24, 24, seed=2)[1]
    if name == "p40":
        return workloads.planted(40, 100, 120, seed=5)[0]
    if name == "p72":
        return workloads.planted(72, 1000, 288, seed=6)[0]
    raise KeyError(name)


# lookahead 0: the plain form's rule (here, with few replicas, 16 groups per
# replica over four waves up to order 32, 8 up to 64), 1: one group per
# replica, 4: four groups in one wave
@pytest.mark.parametrize("look", [0, 1, 4], ids=["auto", "K1", "K4"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_trajectory_matches_model(eng, orc, case, look):
    name, mat, kw = case
    A0 = _matrix(mat)
    n = A0.shape[0]
    epochs = 3
    cap = 200000
    ep, op = _both(max_epochs=epochs, move_log=True, move_log_cap=cap, **kw)
    ep.lookahead = look
    say(f"trajectory {name} n={n} R={ep.replicas} lookahead={look}")
    res = eng.run_annealer(A0, None, ep)
    model = orc.model_run(A0, np.eye(n, dtype=np.int64), op, epochs, log_cap=cap)
    st = res.stats
    R = ep.replicas
    gA, gXt, gsc = eng.last_replicas(R, n)
    assert st["solved"] == model["solved"]
    assert st["verify_status"] == 0
    eye = np.eye(n, dtype=np.int64)
    if model["solved"]:
        r = model["solve_replica"]
        assert st["solve_replica"] == r and st["solve_move"] == model["solve_move"]
        assert np.array_equal(gA[r], model["A"][r]) and np.array_equal(gXt[r], model["Xt"][r])
        assert np.array_equal(res.A, model["A"][r]) and np.array_equal(res.X, model["Xt"][r].T)
        reps = [r]
    else:
        assert st["epochs"] == model["epochs"] == epochs
        assert np.array_equal(gA, model["A"]), "replica A differ"
        assert np.array_equal(gXt, model["Xt"]), "replica Xt differ"
        assert np.array_equal(gsc, model["scal"]), (gsc, model["scal"])
        b = model["best_replica"]
        assert st["best_replica"] == b and st["best_score"] == model["best_score"]
        assert np.array_equal(res.A, model["best_A"][b])
        assert np.array_equal(res.X, model["best_Xt"][b].T)
        assert st["exchanges_accepted"] == model["exchanges"]
        reps = range(R)
    for r in reps:
        glog, trunc = eng.movelog(r)
        assert not trunc
        mlog = model["logs"][r]
        assert len(glog) == len(mlog), (r, len(glog), len(mlog))
        assert np.array_equal(glog, mlog), r
        if (glog["flags"] & 2).any():
            continue  # a reseed jumps to another replica's state: no single-log replay
        k, A, Xt = orc.replay(A0, eye, glog)
        assert k == len(glog)
        assert np.array_equal(A, gA[r]) and np.array_equal(Xt, gXt[r])
    st_o, sign, _ = orc.verify(A0, res.X, res.A)
    assert st_o & ~2 == 0 and sign in (1, -1)


@pytest.mark.parametrize("look", [0, 1, 4], ids=["auto", "K1", "K4"])
def test_trajectory_active_rows_start_transform_and_cap(eng, orc, look):
    """Moves restricted to an active list (as --gram-rows and the deflation
    loop use it), a start transform X0 and a per-replica move cap that falls
    inside an epoch (and inside a lookahead step): every replica ends in the
    oracle model's state, with its accepted-move log, bit for bit."""
    n = 20
    A0 = workloads.random_symmetric(n, -20, 20, seed=21)
    X0 = np.eye(n, dtype=np.int64)
    X0[0, 3], X0[5, 1], X0[7, 12] = 2, -1, 1
    act = [0, 1, 3, 5, 6, 9, 12, 14, 17, 19]
    A_start = (X0.T.astype(object) @ A0.astype(object) @ X0.astype(object)).astype(np.int64)
    cap = 100000
    ep, op = _both(replicas=5, seed=23, exchange_interval=300, t_init=40, t_min=0.5,
                   max_moves=1237, sweep_fraction=0.3, stuck_threshold=150, max_epochs=6,
                   move_log=True, move_log_cap=cap)
    ep.lookahead = look
    say(f"active rows, X0, cap: lookahead={look}")
    res = eng.run_annealer(A0, X0, ep, active=act)
    model = orc.model_run(A_start, X0, op, 6, active=act, log_cap=cap)
    st = res.stats
    gA, gXt, gsc = eng.last_replicas(ep.replicas, n)
    assert st["solved"] == model["solved"] == 0
    assert st["epochs"] == model["epochs"]
    assert np.array_equal(gA, model["A"]) and np.array_equal(gXt, model["Xt"])
    assert np.array_equal(gsc, model["scal"]), (gsc, model["scal"])
    assert (gsc[:, 2] == 1237).all()  # every replica stopped at the cap
    b = model["best_replica"]
    assert np.array_equal(res.A, model["best_A"][b]) and np.array_equal(res.X, model["best_Xt"][b].T)
    for r in range(ep.replicas):
        glog, trunc = eng.movelog(r)
        assert not trunc and np.array_equal(glog, model["logs"][r]), r
        assert set(glog["i"]) <= set(act) and set(glog["j"]) <= set(act)
    # columns outside the active list keep X0's
    for k in range(n):
        if k not in act:
            assert np.array_equal(res.X[:, k], X0[:, k])


# --------------------------------------------------------------- result forms
def test_plain_solves_reference_sample(eng, orc):
    A0 = _matrix("ten")
    ep = EngineParams(replicas=32, seed=1, t_init=50, t_min=0.5, exchange_interval=500,
                      max_epochs=400)
    res = eng.run_annealer(A0, None, ep)
    assert res.stats["solved"] == 1 and res.stats["verify_status"] == 0
    assert sorted(np.diag(res.A).tolist()) == [-1] + [1] * 9
    assert orc.verify(A0, res.X, res.A)[0] == 0


def test_alone_vs_batch_vs_serve(eng):
    mats = [workloads.planted(n, 9, 2 * n, seed=100 + n)[0] for n in (5, 6, 8, 9, 12, 20, 33)]
    mats.append(_matrix("ten"))
    ep = EngineParams(replicas=8, t_init=50, t_min=0.5, exchange_interval=200, max_epochs=50)
    seeds = [1000 + q for q in range(len(mats))]
    alone = []
    for q, A0 in enumerate(mats):
        say(f"alone {q} n={A0.shape[0]}")
        p = EngineParams(**{**ep.__dict__, "seed": seeds[q]})
        r = eng.run_annealer(A0, None, p)
        alone.append((r.A, r.X, r.stats["solved"], r.stats["solve_move"]))
    say("batch")
    batch, bst = eng.batch_solve(mats, ep, seeds=seeds)
    for q in range(len(mats)):
        A, X, d = batch[q]
        assert d["verify_status"] == 0
        assert d["solved"] == alone[q][2]
        if d["solved"]:
            assert d["solve_move"] == alone[q][3]
        assert np.array_equal(A, alone[q][0]) and np.array_equal(X, alone[q][1]), q
    # the rule gives these few replicas lookahead; one group per replica agrees
    say("batch, one group per replica")
    b1, _ = eng.batch_solve(mats, EngineParams(**{**ep.__dict__, "lookahead": 1}), seeds=seeds)
    for q in range(len(mats)):
        A, X, d = b1[q]
        assert d["solved"] == alone[q][2] and d["verify_status"] == 0
        assert np.array_equal(A, alone[q][0]) and np.array_equal(X, alone[q][1]), q
    say("serve")
    eng.serve_open(ep)
    try:
        for q, A0 in enumerate(mats):
            eng.serve_submit(A0, q, seeds[q])
            eng.serve_step()
        got = {}
        for _ in range(2000):
            r, qd = eng.serve_step()
            for A, X, d in eng.serve_poll():
                got[d["tag"]] = (A, X, d)
            if r == 0 and qd == 0:
                break
        for A, X, d in eng.serve_poll():
            got[d["tag"]] = (A, X, d)
        assert len(got) == len(mats)
        for q in range(len(mats)):
            A, X, d = got[q]
            assert d["solved"] == alone[q][2] and d["verify_status"] == 0
            if alone[q][2]:
                assert d["solve_move"] == alone[q][3]
            assert np.array_equal(A, alone[q][0]) and np.array_equal(X, alone[q][1]), q
    finally:
        eng.serve_close()


# ------------------------------------------------------- launch configurations
def _configs(order):
    """(lanes, cols, triangular, waves per workgroup, lookahead)"""
    for G in (16, 32, 64):
        C = -(-order // G)
        if C > 2 or (G, C) == (16, 2):
            continue
        for tri in (1, 0):
            for waves in (1, 2, 4):
                yield G, C, tri, waves, 1
    for K in (2, 4, 8, 16, 32):
        for G in (16, 32):
            C = -(-order // G)
            if C <= 2:
                yield G, C, 1, 0, K
    if order <= 48:  # the lane kernel: one lane per proposal
        for waves in (1, 2, 4):
            yield 1, order, 1, waves, 16


@pytest.mark.parametrize("n", [6, 12, 20, 33, 47, 70])
def test_every_launch_configuration_gives_the_same_result(eng, n):
    """Lanes per replica, columns per lane, LDS layout, waves per workgroup and
    lookahead change speed only: the batch result of every configuration of an order
    equals the rule's, bit for bit (A, X, solve move, best score, and the
    proposal count of an unsolved instance; after a solve the other replicas
    stop at their next poll of the solve key, so that count is timing's)."""
    order = -(-n // 8) * 8
    mats = [workloads.planted(n, 20, 2 * n, seed=300 + q)[0] for q in range(3)]
    mats.append(workloads.random_symmetric(n, -20, 20, seed=7))
    base = dict(replicas=6, t_init=30, t_min=0.5, exchange_interval=300, max_epochs=6)
    ep = EngineParams(**base)
    seeds = [77 + q for q in range(len(mats))]
    eng.launch_reset(order)
    ref, _ = eng.batch_solve(mats, ep, seeds=seeds)
    tried = 0
    try:
        for G, C, tri, waves, K in _configs(order):
            try:
                eng.launch_set(order, G, C, tri, waves, K)
            except RuntimeError:
                continue  # no kernel with that shape, or its LDS does not fit
            cfg = eng.launch_get(n)
            assert (cfg["lanes"], cfg["cols"], cfg["triangular"], cfg["tuned"],
                    cfg["lookahead"]) == (G, C, tri, 1, K)
            say(f"n={n} G={G} C={C} tri={tri} waves={waves} lookahead={K}")
            # these few replicas get a lookahead layout by the rule: a one-group
            # configuration applies where one group is requested, a lookahead
            # one under the rule
            got, _ = eng.batch_solve(mats, EngineParams(**base, lookahead=1 if K == 1 else 0),
                                     seeds=seeds)
            tried += 1
            for q in range(len(mats)):
                A, X, d = got[q]
                assert d["verify_status"] == 0
                assert np.array_equal(A, ref[q][0]) and np.array_equal(X, ref[q][1]), (q, G, C)
                keys = ("solved", "solve_move", "best_score") + (() if d["solved"] else ("moves",))
                for k in keys:
                    assert d[k] == ref[q][2][k], (k, q, G, C, tri, waves, K)
    finally:
        eng.launch_reset(order)
    assert tried >= 2


def test_lookahead_changes_speed_only(eng):
    """Evaluating a replica's next proposals at once (lookahead 2 or 4: one
    wave per replica; 8, 16 or 32: a workgroup of four or eight waves; the lane groups
    resolve the proposals in move order) gives the batch result of one group
    per replica, bit for bit, at every order (above 64 the engine keeps one
    group), with stall sweeps among the moves; a lookahead other than 0, 1, 2,
    4, 8, 16 or 32 is refused."""
    from xtax_amd._native import XtxError
    mats = [workloads.planted(n, 20, 2 * n, seed=500 + n)[0] for n in (5, 12, 24, 31, 40, 60, 70)]
    mats.append(workloads.random_symmetric(16, -20, 20, seed=8))
    seeds = [900 + q for q in range(len(mats))]
    base = dict(replicas=6, t_init=30, t_min=0.5, exchange_interval=300, max_epochs=6,
                sweep_fraction=0.3, stuck_threshold=200)
    ref, _ = eng.batch_solve(mats, EngineParams(lookahead=1, **base), seeds=seeds)
    for look in (2, 4, 8, 16, 32):
        say(f"batch lookahead={look}")
        got, _ = eng.batch_solve(mats, EngineParams(lookahead=look, **base), seeds=seeds)
        for q in range(len(mats)):
            A, X, d = got[q]
            assert d["verify_status"] == 0
            assert np.array_equal(A, ref[q][0]) and np.array_equal(X, ref[q][1]), (q, look)
            keys = ("solved", "solve_move", "best_score") + (() if d["solved"] else ("moves",))
            for k in keys:
                assert d[k] == ref[q][2][k], (k, q, look)
    with pytest.raises(XtxError):
        eng.run_annealer(mats[0], None, EngineParams(lookahead=3, replicas=2, max_epochs=1))


def test_autotune_picks_a_measured_configuration(eng):
    mats = [workloads.planted(n, 20, 2 * n, seed=400 + n)[0] for n in (7, 14, 22, 40)]
    ep = EngineParams(replicas=4, t_init=30, t_min=0.5, exchange_interval=200, max_epochs=3)
    seeds = [5 + q for q in range(len(mats))]
    ref, _ = eng.batch_solve(mats, ep, seeds=seeds)
    try:
        res = eng.autotune(mats, ep, moves=64, seeds=seeds)
        for n in (7, 14, 22, 40):
            order = -(-n // 8) * 8
            rows = [r for r in res if r["order"] == order]
            assert rows and all(r["moves_per_sec"] > 0 for r in rows)
            chosen = [r for r in rows if r["tuned"]]
            assert len(chosen) == 1
            assert chosen[0]["moves_per_sec"] == max(r["moves_per_sec"] for r in rows)
            cfg = eng.launch_get(n)
            assert cfg["tuned"] == 1
            assert (cfg["lanes"], cfg["cols"], cfg["triangular"], cfg["waves_per_block"],
                    cfg["lookahead"]) == (
                chosen[0]["lanes"], chosen[0]["cols"], chosen[0]["triangular"],
                chosen[0]["waves_per_block"], chosen[0]["lookahead"])
        got, _ = eng.batch_solve(mats, ep, seeds=seeds)
        for q in range(len(mats)):
            assert np.array_equal(got[q][0], ref[q][0]) and np.array_equal(got[q][1], ref[q][1])
    finally:
        for order in (8, 16, 24, 40):
            eng.launch_reset(order)


def test_new_best_reports(eng, orc):
    """on_best: one report per epoch whose instance best improved, strictly
    decreasing, each equal to the best-so-far score the replicas record."""
    import ctypes as C
    from xtax_amd import _native as N
    A0 = workloads.planted(24, 100, 72, seed=3)[0]
    seen = []
    snaps = []

    def on_best(user, t, score, dscore, rep):
        seen.append((t, score, dscore, rep))
        A1 = np.zeros((n, n), np.int64)
        X1 = np.zeros((n, n), np.int64)
        assert N.lib().xtx_best_so_far(eng._h, A1.ctypes.data_as(N.I64P), None,
                                       X1.ctypes.data_as(N.I64P)) == 0
        snaps.append((score, A1, X1))

    n = A0.shape[0]
    cb = N.ON_BEST(on_best)
    ep = EngineParams(replicas=16, seed=4, max_epochs=12, exchange_interval=500, t_init=18.0,
                      t_min=14.0)
    p = ep.to_c()
    p.on_best = C.cast(cb, C.c_void_p)
    A = np.zeros((n, n), np.int64)
    X = np.zeros((n, n), np.int64)
    st = N.Stats()
    A0c = np.ascontiguousarray(A0, np.int64)
    rc = N.lib().xtx_solve_plain(eng._h, A0c.ctypes.data_as(N.I64P), None, n, C.byref(p), None,
                                 0, A.ctypes.data_as(N.I64P), X.ctypes.data_as(N.I64P),
                                 C.byref(st))
    assert rc == 0
    start = int(2 * np.abs(A0).sum() - np.abs(np.diag(A0)).sum())
    assert seen, "no report"
    scores = [start] + [s for _, s, _, _ in seen]
    assert all(b < a for a, b in zip(scores, scores[1:]))
    assert all(float(s) == d for _, s, d, _ in seen)
    assert all(0 <= r < 16 for _, _, _, r in seen)
    assert len(seen) <= st.epochs
    if not st.solved:
        assert seen[-1][1] == st.best_score
    # each report's snapshot is that best: X^T A0 X = A exactly, with its score
    for score, A1, X1 in snaps:
        status, sign, _ = orc.verify(A0, X1, A1)
        assert status & ~2 == 0
        assert int(2 * np.abs(A1).sum() - np.abs(np.diag(A1)).sum()) == score
    assert N.lib().xtx_best_so_far(eng._h, None, None, None) != 0   # outside a report
