"""The lane kernel (one lane per proposal, 16 or 32 proposals per replica and
step, four or two replicas per wave; two orders per launch; anneal_lane.hip)
against the oracle's trajectory model and against the lane-group kernels, bit
for bit.

1. Trajectories: plain solves of three epochs with the lane configuration set
   for the order (xtx_launch_set: lanes 1, cols = order, lookahead 16 or 32) end in
   the model's state -- every replica's A, X^T and scalars, the returned best
   and the accepted-move log (which replays to the final state) -- over the
   matrix families, both temperature modes with stall sweeps, reseeds,
   adaptive cooling, every tournament size, an active list with a start
   transform and a move cap inside a step, rows past 2^16 (the int64 path),
   and waves per workgroup 0, 1, 2.
2. 24 randomly drawn configurations at orders up to 48.
3. Batches: the same instances under the lane kernel (orders paired in one
   launch, both orders present in every launch) and under the lookahead
   kernels give the same result for every instance, at 16 and 32 proposals
   per step.
"""
import sys

import numpy as np
import pytest

from xtax_amd import EngineParams, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _lane(eng, n, waves=0, q=16):
    order = max(8, -(-n // 8) * 8)
    eng.launch_set(order, 1, order, 1, waves, q)
    cfg = eng.launch_get(n)
    assert (cfg["lanes"], cfg["cols"], cfg["lookahead"], cfg["tuned"]) == (1, order, q, 1)
    return order


def _trajectory(eng, orc, A0, kw, epochs=3, X0=None, active=None, log_cap=100000):
    n = A0.shape[0]
    from oracle import pyoracle as po
    ep = EngineParams(max_epochs=epochs, move_log=True, move_log_cap=log_cap, **kw)
    op = po.make_params(**{a: b for a, b in kw.items() if a in po.DEFAULTS})
    res = eng.run_annealer(A0, X0, ep, active=active)
    Xs = np.eye(n, dtype=np.int64) if X0 is None else X0
    A_start = A0 if X0 is None else (Xs.T.astype(object) @ A0.astype(object)
                                     @ Xs.astype(object)).astype(np.int64)
    m = po.model_run(A_start, Xs, op, epochs, active=active, log_cap=log_cap)
    st = res.stats
    R = kw["replicas"]
    gA, gXt, gsc = eng.last_replicas(R, n)
    assert st["solved"] == m["solved"] and st["verify_status"] == 0
    if m["solved"]:
        r = m["solve_replica"]
        assert st["solve_replica"] == r and st["solve_move"] == m["solve_move"]
        assert np.array_equal(gA[r], m["A"][r]) and np.array_equal(gXt[r], m["Xt"][r])
        reps = [r]
    else:
        assert st["epochs"] == m["epochs"]
        assert np.array_equal(gA, m["A"]), "replica A differ"
        assert np.array_equal(gXt, m["Xt"]), "replica Xt differ"
        assert np.array_equal(gsc, m["scal"]), (gsc, m["scal"])
        b = m["best_replica"]
        assert np.array_equal(res.A, m["best_A"][b])
        assert np.array_equal(res.X, m["best_Xt"][b].T)
        assert st["exchanges_accepted"] == m["exchanges"]
        reps = range(R)
    for r in reps:
        glog, trunc = eng.movelog(r)
        assert not trunc and np.array_equal(glog, m["logs"][r]), r
        if X0 is None and active is None and not (glog["flags"] & 2).any():
            k, A, Xt = orc.replay(A0, np.eye(n, dtype=np.int64), glog)
            assert k == len(glog) and np.array_equal(A, gA[r]) and np.array_equal(Xt, gXt[r])
    return res, m


CASES = [
    ("10x10-temper", lambda: np.load("tests/golden/g2_end_to_end.npz")["r0_A0"],
     dict(replicas=8, seed=3, exchange_interval=300, t_init=50, t_min=0.5)),
    ("10x10-cool", lambda: np.load("tests/golden/g2_end_to_end.npz")["r0_A0"],
     dict(replicas=4, seed=5, tempering=False, exchange_interval=200, stuck_threshold=150,
          sweep_fraction=0.5, adaptive_cooling=True, moves_per_cool=50)),
    ("rand20-sweeps", lambda: workloads.random_symmetric(20, -20, 20, seed=4),
     dict(replicas=6, seed=9, exchange_interval=250, t_init=40, t_min=0.5, target_samples=3,
          stuck_threshold=100, sweep_fraction=0.5)),
    ("lattice24", lambda: workloads.lattice(24, 24, seed=2)[1],
     dict(replicas=5, seed=11, exchange_interval=200, greedy_fraction=0.5, sweep_fraction=0.05,
          stuck_threshold=120, target_samples=16)),
    ("planted40-greedy", lambda: workloads.planted(40, 100, 120, seed=5)[0],
     dict(replicas=4, seed=13, exchange_interval=150, t_init=200, t_min=5, greedy_fraction=1.0,
          worker_diversity=0.0)),
    ("planted48-big", lambda: workloads.planted(48, 1000, 192, seed=6)[0],
     dict(replicas=3, seed=17, exchange_interval=120, t_init=300, t_min=5, greedy_fraction=0.7)),
    ("rand12-cool-reseed", lambda: workloads.random_symmetric(12, -20, 20, seed=6),
     dict(replicas=3, seed=19, tempering=False, exchange_interval=100, stuck_threshold=60,
          reseed_factor=1.0, t_init=30)),
    ("rand8-uniform", lambda: workloads.random_symmetric(8, -20, 20, seed=8),
     dict(replicas=8, seed=23, exchange_interval=200, t_init=20, t_min=1, target_fraction=0.0,
          greedy_fraction=0.5)),
    ("bound-edge", lambda: _edge(16),
     dict(replicas=4, seed=29, exchange_interval=100, t_init=1e12, t_min=1e6)),
]


def _edge(n):
    rng = np.random.default_rng(31)
    A = rng.integers(-(1 << 47), (1 << 47) + 1, (n, n)) * (rng.random((n, n)) < 0.6)
    A = np.triu(A) + np.triu(A, 1).T
    A[np.arange(n), np.arange(n)] = rng.integers(1, 1 << 20, n)
    return A.astype(np.int64)


@pytest.mark.parametrize("q", [16, 32])
@pytest.mark.parametrize("waves", [0, 2])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lane_trajectory_matches_model(eng, orc, case, waves, q):
    name, mk, kw = case
    A0 = mk()
    n = A0.shape[0]
    order = _lane(eng, n, waves, q)
    try:
        say(f"lane trajectory {name} n={n} R={kw['replicas']} waves={waves} q={q}")
        _trajectory(eng, orc, A0, kw)
    finally:
        eng.launch_reset(order)


def test_lane_active_rows_start_transform_and_cap(eng, orc):
    n = 20
    A0 = workloads.random_symmetric(n, -20, 20, seed=21)
    X0 = np.eye(n, dtype=np.int64)
    X0[0, 3], X0[5, 1], X0[7, 12] = 2, -1, 1
    act = [0, 1, 3, 5, 6, 9, 12, 14, 17, 19]
    order = _lane(eng, n)
    try:
        say("lane: active rows, X0, cap 1237")
        res, m = _trajectory(eng, orc, A0, dict(replicas=5, seed=23, exchange_interval=300,
                                                t_init=40, t_min=0.5, max_moves=1237,
                                                sweep_fraction=0.3, stuck_threshold=150),
                             epochs=6, X0=X0, active=act)
        gA, gXt, gsc = eng.last_replicas(5, n)
        assert (gsc[:, 2] == 1237).all()
    finally:
        eng.launch_reset(order)


def _draw(k):
    rng = np.random.default_rng(5000 + k)
    n = int(rng.choice([3, 5, 8, 11, 16, 17, 24, 31, 33, 40, 47, 48]))
    fam = int(rng.integers(0, 3))
    if fam == 0:
        A0 = workloads.random_symmetric(n, -20, 20, seed=k)
    elif fam == 1:
        A0 = workloads.planted(n, 100, 3 * n, seed=k)[0]
    else:
        A0 = workloads.lattice(n, n + 2, seed=k)[1]
    kw = dict(replicas=int(rng.choice([1, 2, 3, 5, 8, 13])), seed=int(rng.integers(0, 2 ** 40)),
              tempering=bool(rng.integers(0, 2)), exchange_interval=int(rng.integers(50, 400)),
              greedy_fraction=float(rng.choice([0.0, 0.3, 0.5, 1.0])),
              target_fraction=float(rng.choice([0.0, 0.5, 1.0])),
              target_samples=int(rng.choice([1, 3, 5, 8, 16])),
              sweep_fraction=float(rng.choice([0.0, 0.2])),
              stuck_threshold=int(rng.choice([50, 300, 20000])),
              worker_diversity=float(rng.choice([0.0, 0.4])),
              t_init=float(rng.choice([0.0, 5.0, 50.0])), t_min=float(rng.choice([0.1, 2.0])),
              reseed_factor=float(rng.choice([1.0, 1.25])),
              adaptive_cooling=bool(rng.integers(0, 2)),
              moves_per_cool=int(rng.choice([1, 20, 200])))
    return A0, kw


@pytest.mark.parametrize("k", range(24))
def test_lane_random_config(eng, orc, k):
    A0, kw = _draw(k)
    n = A0.shape[0]
    if np.count_nonzero(np.triu(A0, 1)) == 0:
        pytest.skip("diagonal start")
    q = 32 if k % 2 else 16
    order = _lane(eng, n, k % 3, q)
    try:
        say(f"lane random config {k} n={n} q={q} {kw}")
        _trajectory(eng, orc, A0, kw)
    finally:
        eng.launch_reset(order)


@pytest.mark.parametrize("count,reps,q", [(96, 32, 0), (384, 16, 0), (160, 32, 16)])
def test_lane_batch_equals_lookahead_batch(eng, count, reps, q):
    """A batch large enough for the rule to choose the lane kernel (32
    proposals per step) gives the result of the same batch under the
    lookahead kernels (explicit lookahead 16), instance by instance. q = 16:
    the lane configuration at 16 proposals is set for every order."""
    mats = workloads.batch_config(count=count, seed=11)
    orders = sorted({max(8, -(-a.shape[0] // 8) * 8) for a in mats})
    assert orders == [8, 16, 24, 32, 40, 48]
    seeds = [3_000_017 + q for q in range(count)]
    base = dict(replicas=reps, t_init=18.0, t_min=14.0, target_fraction=1.0,
                exchange_interval=1000, max_epochs=3)
    say(f"lane batch {count} x {reps} q={q}")
    if q:
        for o in orders:
            eng.launch_set(o, 1, o, 1, 0, q)
    try:
        lane, bl = eng.batch_solve(mats, EngineParams(**base), seeds=seeds)
    finally:
        for o in orders:
            eng.launch_reset(o)
    if q:
        assert bl["lane_instances"] == count
    else:
        assert bl["lane_instances"] == (count if count * reps >= 4096 else 0)
    grp, bg = eng.batch_solve(mats, EngineParams(lookahead=16, **base), seeds=seeds)
    say(f"lane {bl['kernel_ms']:.1f} ms, lookahead {bg['kernel_ms']:.1f} ms, solved {bl['solved']}")
    assert bl["solved"] == bl["verified"]
    for q in range(count):
        A, X, d = lane[q]
        A2, X2, d2 = grp[q]
        assert d["verify_status"] == 0 and d["solved"] == d2["solved"], q
        keys = ("solve_move", "best_score") + (() if d["solved"] else ("moves",))
        for k in keys:
            assert d[k] == d2[k], (k, q)
        assert np.array_equal(A, A2) and np.array_equal(X, X2), q
