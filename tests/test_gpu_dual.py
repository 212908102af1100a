"""Parity of the engine on xdual's primal/dual objective (anneal_dual.hip,
dual.cpp, through the C ABI) with the oracle (oracle/oracle_dual.cpp).

1. Primitives: the device's move evaluation (feasibility, nonzero change,
   delta d_off_p + c d_off_q, proposed P and Q rows), best shear and score are
   the oracle's bit for bit, on the reference's own matrices (G4 golden) and on
   random states for every lane-group layout (n <= 16, 32, 64, 128).
2. Trajectories: a solve for a fixed number of epochs ends in exactly the
   state of the oracle's trajectory model (every replica's P, Q, X^T, score,
   best, local best, temperature, counters), with the maintenance (dual
   re-inversion, residual checks, rescoring) exercised; move logs replay P.
3. Results: solved instances pass the exact verifier with Q = P^-1 to
   rounding; a singular Gram is refused.
"""
import os
import sys

import numpy as np
import pytest

from xtax_amd import DualParams, EngineParams

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "g4_xdual.npz")


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def po():
    from oracle import pyoracle
    return pyoracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gram(n, seed, spread=3, mults=3):
    """B diag(d) B^T with B a product of random integer shears (nonsingular)."""
    rng = np.random.default_rng(seed)
    B = np.eye(n, dtype=np.int64)
    for _ in range(mults * n):
        a, b = rng.choice(n, 2, replace=False)
        B[a] += int(rng.integers(-spread + 1, spread)) * B[b]
    d = rng.integers(1, 4, n)
    return (B @ np.diag(d) @ B.T).astype(np.int64)


def _check(eng, po, P, Q, c, queries):
    oi, od, rp, rq, sc = eng.probe_dual(P, Q, c, queries)
    n = P.shape[-1]
    for q, (m, i, j, s) in enumerate(queries):
        Pm, Qm = (P[m], Q[m]) if P.ndim == 3 else (P, Q)
        ok, dp, dq, dnz, prow, qrow = po.dual_eval(Pm, Qm, i, j, s)
        assert bool(oi[q, 0]) == ok, (n, q, i, j, s)
        assert oi[q, 2] == po.dual_best_shear(Pm, Qm, c, i, j), (n, q, i, j)
        if ok:
            assert oi[q, 1] == dnz
            want = dp + c * dq
            assert od[q] == want, (n, q, od[q], want)       # bit for bit
            assert np.array_equal(rp[q], prow)
            assert np.array_equal(rq[q], qrow), (n, q)
        assert sc[q] == po.dual_score(Pm, Qm, c)


def test_primitives_on_reference_matrices(eng, po):
    g4 = np.load(GOLD)
    for t in range(int(g4["cases"])):
        P, Q, c = g4[f"c{t}_P"], g4[f"c{t}_Q"], float(g4[f"c{t}_c"])
        qs = [(0, int(i), int(j), int(s)) for i, j, s in g4[f"c{t}_q"]]
        say("g4 case", t, "n", P.shape[0])
        _check(eng, po, P, Q, c, qs)


@pytest.mark.parametrize("n", [5, 16, 17, 32, 40, 64, 65, 100, 128])
def test_primitives_every_layout(eng, po, n):
    rng = np.random.default_rng(n)
    P = gram(n, n)
    ok, Q = po.invert(P)
    assert ok
    c = po.frob2_off_p(P) / po.frob2_off_q(Q)
    qs = []
    for _ in range(200):
        i, j = (int(v) for v in rng.choice(n, 2, replace=False))
        s = int(rng.choice([-64, -3, -1, 1, 2, 7, 64, int(rng.integers(-64, 65))]))
        qs.append((0, i, j, s))
    say("layout n", n)
    _check(eng, po, P, Q, c, qs)
    # a state with entries near the bound: infeasible proposals too
    big = P.copy()
    big[0, 1] = big[1, 0] = (1 << 48) - 5
    _check(eng, po, big, Q, c, [(0, 0, 1, 1), (0, 1, 0, -1), (0, 2, 1, 3), (0, 1, 2, 64)])


TRAJ = [  # n, R, tempering, extra engine settings, dual_refresh, dual_check
    (9, 4, 1, {}, 700, 250),
    (9, 4, 0, {}, 700, 250),
    (20, 8, 1, {}, 700, 250),
    (40, 4, 0, {}, 700, 250),
    (70, 2, 1, {}, 700, 250),
    # cold replicas that sit on their best while the dual is re-inverted
    # every epoch: the best keeps the dual it was found with
    # (both end unsolved on a best found before the last re-inversion)
    (20, 4, 0, dict(t_init=0.05, t_min=0.01, stuck_threshold=100000), 300, 0),
    (20, 4, 1, dict(t_init=0.5, t_min=0.1, stuck_threshold=100000), 300, 0),
]


# lookahead 0: the rule (here 16, 8 or 4 lane groups per replica over a
# four-wave workgroup for n <= 64), 1: one group, 2/4: one wave's groups
@pytest.mark.parametrize("look", [0, 1, 4], ids=["auto", "K1", "K4"])
@pytest.mark.parametrize("n,R,tempering,extra,refresh,check", TRAJ)
def test_trajectory_equals_model(eng, po, n, R, tempering, extra, refresh, check, look):
    seed = 100 + n
    A0 = gram(n, seed)
    kw = dict(replicas=R, seed=seed, tempering=tempering, exchange_interval=300,
              stuck_threshold=150, sweep_fraction=0.5, greedy_fraction=0.6)
    kw.update(extra)
    ep = EngineParams(max_epochs=6 if extra else 4, move_log=True, move_log_cap=4096,
                      verify=True, lookahead=look, **kw)
    dpar = DualParams(dual_refresh=refresh, dual_check=check, dual_tol=1e-9)
    say("dual trajectory n", n, "R", R, "tempering", tempering, extra, "lookahead", look)
    res = eng.run_xdual(A0, None, ep, dpar, allow_unverified=False)
    st = res.stats
    P, Q, Xt, dsc, sc = eng.dual_last_replicas(R, n)
    ok, Q0 = po.invert(A0)
    assert ok
    mp = po.Params(**dict(po.DEFAULTS, **kw))
    md = po.DualParams(c=st["c"], dual_tol=1e-9, dual_refresh=refresh, dual_check=check)
    m = po.dual_model_run(A0, Q0, np.eye(n, dtype=np.int64), mp, md, st["epochs"], log_cap=4096)
    assert st["c"] == pytest.approx(po.frob2_off_p(A0) / po.frob2_off_q(Q0), rel=1e-12)
    assert st["solved"] == m["solved"]
    assert st["verify_status"] == 0
    if m["solved"]:
        # after a solve the other replicas stop at a poll point of their own
        reps = [m["solve_replica"]]
        assert st["solve_replica"] == reps[0] and st["solve_move"] == m["solve_move"]
    else:
        reps = range(R)
        assert st["epochs"] == m["epochs"]
        assert st["refreshes"] == m["refreshes"] and st["checks"] == m["checks"]
        assert st["moves"] == m["total_moves"]
        assert st["best_replica"] == m["best_replica"]
    eye = np.eye(n, dtype=np.int64)
    for r in reps:
        assert np.array_equal(P[r], m["P"][r]), r
        assert np.array_equal(Xt[r], m["Xt"][r]), r
        assert np.array_equal(Q[r], m["Q"][r]), r
        assert dsc[r].tolist() == m["dscal"][r].tolist(), (r, dsc[r], m["dscal"][r])
        assert sc[r].tolist() == m["scal"][r].tolist(), (r, sc[r], m["scal"][r])
        log, trunc = eng.movelog(r)
        ml = m["logs"][r]
        assert not trunc and len(log) == len(ml), (r, len(log), len(ml))
        assert np.array_equal(log.view(np.uint8), ml.view(np.uint8)), r
        if (log["flags"] & 2).any():
            continue  # a reseed jumps to the instance best: no single-log replay
        k, Pr, Xr = po.replay(A0, eye, log)   # the primal moves are xtax's Add shears
        assert k == len(log)
        assert np.array_equal(Pr, P[r]) and np.array_equal(Xr, Xt[r])
    assert np.array_equal(res.P, m["best_P"]) and np.array_equal(res.Q, m["best_Q"])
    assert np.array_equal(res.X.T, m["best_Xt"])
    if extra:
        b = m["best_replica"]
        assert not m["solved"] and np.array_equal(m["best_P"], m["P"][b])
        assert not np.array_equal(m["best_Q"], m["Q"][b])  # re-inverted since


def test_solves_and_verifies(eng, po):
    n = 12
    A0 = gram(n, 7, spread=2, mults=2)
    ep = EngineParams(replicas=16, seed=3, max_seconds=20.0, verify=True)
    res = eng.run_xdual(A0, None, ep, DualParams())
    st = res.stats
    say("dual solve", {k: st[k] for k in ("solved", "moves", "best_score", "epochs", "seconds")})
    assert st["solved"] == 1 and st["verify_status"] == 0
    P, Q, X = res.P, res.Q, res.X
    assert np.count_nonzero(P - np.diag(np.diag(P))) == 0
    Xo = X.astype(object)
    assert np.array_equal((Xo.T @ A0.astype(object) @ Xo).astype(np.int64), P)
    assert np.abs(P.astype(float) @ Q - np.eye(n)).max() < 1e-6
    assert st["best_score"] == pytest.approx(po.dual_score(P, Q, st["c"]), rel=1e-12)


def test_start_transform_and_active_rows(eng, po):
    n = 10
    A0 = gram(n, 11)
    X0 = np.eye(n, dtype=np.int64)
    X0[0, 3] = 2
    X0[5, 1] = -1
    act = [0, 2, 3, 5, 7, 8]
    ep = EngineParams(replicas=4, seed=9, max_epochs=3, exchange_interval=200, verify=True)
    res = eng.run_xdual(A0, X0, ep, DualParams(), active=act)
    Xo = res.X.astype(object)
    assert np.array_equal((Xo.T @ A0.astype(object) @ Xo).astype(np.int64), res.P)
    # rows outside the active set never move: X keeps X0's columns there
    for k in range(n):
        if k not in act:
            assert np.array_equal(res.X[:, k], X0[:, k])


def test_singular_gram_is_refused(eng):
    from xtax_amd._native import XtxError
    # rank 2, and exactly so in floating point: the second pivot is 0.0 (a
    # singular matrix whose elimination rounds to a tiny pivot is inverted,
    # as invert_to does)
    A0 = np.array([[1, 2, 0], [2, 4, 0], [0, 0, 1]], np.int64)
    with pytest.raises(XtxError, match="invert"):
        eng.run_xdual(A0, None, EngineParams(replicas=2, max_epochs=1), DualParams())


def test_limits_and_small_orders(eng, po):
    """max_moves caps every replica, max_seconds stops the run, n = 2 and a
    two-index active list work, an already diagonal start is reported solved."""
    A0 = gram(14, 31)
    r = eng.run_xdual(A0, None, EngineParams(replicas=4, seed=2, max_moves=700,
                                             exchange_interval=300), DualParams())
    assert r.stats["stop_reason"] == 3 and r.stats["moves"] == 4 * 700
    r = eng.run_xdual(A0, None, EngineParams(replicas=8, seed=2, max_seconds=0.3,
                                             exchange_interval=2000), DualParams())
    assert r.stats["stop_reason"] in (0, 1) and r.stats["verify_status"] == 0
    A2 = np.array([[2, 2], [2, 5]], np.int64)   # U^T diag(2, 3) U, U = [[1, 1], [0, 1]]
    r = eng.run_xdual(A2, None, EngineParams(replicas=2, seed=1, max_epochs=20), DualParams())
    assert r.stats["solved"] == 1 and r.P[0, 1] == 0
    assert np.array_equal(r.X.T @ A2 @ r.X, r.P)
    r = eng.run_xdual(A0, None, EngineParams(replicas=2, seed=1, max_epochs=3), DualParams(),
                      active=[3, 9])
    keep = [k for k in range(14) if k not in (3, 9)]
    assert np.array_equal(r.X[:, keep], np.eye(14, dtype=np.int64)[:, keep])
    D = np.diag([3, -1, 2]).astype(np.int64)
    r = eng.run_xdual(D, None, EngineParams(replicas=2), DualParams())
    assert r.stats["solved"] == 1 and r.stats["stop_reason"] == 4
    assert np.allclose(r.Q, np.diag([1 / 3, -1, 0.5]))


def test_new_best_reports(eng, po):
    """on_best reports the xdual instance best (dscore) as it improves."""
    import ctypes as C
    from xtax_amd import _native as N
    n = 20
    A0 = gram(n, 120)
    seen = []
    snaps = []

    def on_best(user, t, score, dscore, rep):
        seen.append((score, dscore, rep))
        P1 = np.zeros((n, n), np.int64)
        Q1 = np.zeros((n, n))
        X1 = np.zeros((n, n), np.int64)
        assert N.lib().xtx_best_so_far(eng._h, P1.ctypes.data_as(N.I64P),
                                       Q1.ctypes.data_as(C.POINTER(C.c_double)),
                                       X1.ctypes.data_as(N.I64P)) == 0
        snaps.append((P1, Q1, X1))

    cb = N.ON_BEST(on_best)
    p = EngineParams(replicas=8, seed=3, max_epochs=8, exchange_interval=300).to_c()
    p.on_best = C.cast(cb, C.c_void_p)
    d = DualParams().to_c()
    P = np.zeros((n, n), np.int64)
    Q = np.zeros((n, n))
    X = np.zeros((n, n), np.int64)
    st = N.DualStats()
    A0c = np.ascontiguousarray(A0, np.int64)
    dp = C.POINTER(C.c_double)
    rc = N.lib().xtx_dual_solve(eng._h, A0c.ctypes.data_as(N.I64P), None, n, C.byref(p),
                                C.byref(d), None, 0, P.ctypes.data_as(N.I64P),
                                Q.ctypes.data_as(dp), X.ctypes.data_as(N.I64P), C.byref(st))
    assert rc == 0 and seen
    scores = [st.start_score] + [ds for _, ds, _ in seen]
    assert all(b < a for a, b in zip(scores, scores[1:]))
    assert all(s == 0 and 0 <= r < 8 for s, _, r in seen)
    for P1, Q1, X1 in snaps:
        Xo = X1.astype(object)
        assert np.array_equal((Xo.T @ A0.astype(object) @ Xo).astype(np.int64), P1)
        assert np.abs(P1.astype(float) @ Q1 - np.eye(n)).max() < 1e-4
