"""Orders above 128 and tournaments above 16 samples (the reference limits
neither): the lane kernel's wide mode (anneal_lane.hip, WIDE: A exact in HBM,
32 proposals per replica and step, each evaluated by the replica's 32 lanes,
tournaments of any size) against the oracle's trajectory model, bit for bit
-- every replica's A, X^T and scalars and the accepted-move log -- at orders
130 to 256, tempering and cooling with stall sweeps and reseeds; tournaments
of 17 to 300 samples at every order (their draws past slot 15 on their own
counters, the model pinned against the reference with such tournaments in
tests/test_model_pinned.py); a batch mixing wide and lane-kernel orders equal
to its instances run alone; the command line end to end at n = 140."""
import os
import subprocess
import sys

import numpy as np
import pytest

from xtax_amd import EngineParams, workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _trajectory(eng, orc, A0, kw, epochs, log_cap=3000):
    from oracle import pyoracle as po
    n = A0.shape[0]
    ep = EngineParams(max_epochs=epochs, move_log=True, move_log_cap=log_cap, **kw)
    op = po.make_params(**{a: b for a, b in kw.items() if a in po.DEFAULTS})
    res = eng.run_annealer(A0, None, ep)
    m = orc.model_run(A0, np.eye(n, dtype=np.int64), op, epochs, log_cap=log_cap)
    st = res.stats
    R = kw["replicas"]
    gA, gXt, gsc = eng.last_replicas(R, n)
    assert st["solved"] == m["solved"] and st["verify_status"] == 0
    if m["solved"]:
        r = m["solve_replica"]
        assert st["solve_move"] == m["solve_move"]
        assert np.array_equal(gA[r], m["A"][r]) and np.array_equal(gXt[r], m["Xt"][r])
        reps = [r]
    else:
        assert np.array_equal(gA, m["A"]) and np.array_equal(gXt, m["Xt"])
        assert np.array_equal(gsc, m["scal"])
        b = m["best_replica"]
        assert np.array_equal(res.A, m["best_A"][b]) and np.array_equal(res.X, m["best_Xt"][b].T)
        assert st["exchanges_accepted"] == m["exchanges"]
        reps = range(R)
    for r in reps:
        glog, trunc = eng.movelog(r)
        assert not trunc and np.array_equal(glog, m["logs"][r]), r
    return res, m


@pytest.mark.parametrize("n", [130, 150, 192, 200, 256])
def test_wide_tempering_matches_model(eng, orc, n):
    A0 = workloads.planted(n, 100, 2 * n, seed=n)[0]
    cfg = eng.launch_get(n)
    assert (cfg["lanes"], cfg["lookahead"], cfg["order"]) == (1, 32, 192 if n <= 192 else 256)
    say(f"wide n={n} tempering")
    _trajectory(eng, orc, A0, dict(replicas=3, seed=n, t_init=200.0, t_min=5.0,
                                   greedy_fraction=0.6, exchange_interval=150), epochs=3)


@pytest.mark.parametrize("n", [140, 250])
def test_wide_cooling_sweeps_reseeds(eng, orc, n):
    A0 = workloads.random_symmetric(n, -20, 20, seed=n)
    say(f"wide n={n} cooling, sweeps, reseeds")
    _trajectory(eng, orc, A0, dict(replicas=2, seed=n + 1, tempering=False, t_init=30.0,
                                   exchange_interval=120, stuck_threshold=40,
                                   sweep_fraction=0.5, reseed_factor=1.0, target_samples=5),
                epochs=3)


@pytest.mark.parametrize("n,ts,temper", [(5, 17, 1), (12, 24, 1), (40, 40, 0), (70, 100, 1),
                                         (100, 33, 0), (150, 300, 1), (256, 20, 0)])
def test_large_tournaments_match_model(eng, orc, n, ts, temper):
    A0 = (workloads.planted(n, 100, 2 * n, seed=n + ts)[0] if n % 2 == 0
          else workloads.random_symmetric(n, -20, 20, seed=n + ts))
    cfg = eng.launch_get(n)
    say(f"target_samples={ts} n={n} tempering={temper}")
    kw = dict(replicas=3, seed=ts, target_samples=ts, target_fraction=0.7, t_init=100.0,
              t_min=2.0, exchange_interval=160, tempering=bool(temper))
    if not temper:
        kw.update(stuck_threshold=50, sweep_fraction=0.5, reseed_factor=1.0)
    _trajectory(eng, orc, A0, kw, epochs=3)


def test_large_tournament_batch_equals_alone(eng):
    mats = [workloads.planted(n, 100, 3 * n, seed=n)[0] for n in (9, 30, 47, 64, 90, 160)]
    seeds = [501 + q for q in range(len(mats))]
    kw = dict(replicas=4, t_init=60.0, t_min=10.0, exchange_interval=200, max_epochs=3,
              target_samples=24, target_fraction=1.0)
    say("batch with 24-sample tournaments")
    res, st = eng.batch_solve(mats, EngineParams(**kw), seeds=seeds)
    for q, A0 in enumerate(mats):
        alone = eng.run_annealer(A0, None, EngineParams(seed=seeds[q], **kw))
        A, X, d = res[q]
        assert d["verify_status"] == 0 and d["solved"] == alone.stats["solved"], q
        assert np.array_equal(A, alone.A) and np.array_equal(X, alone.X), q


def test_xdual_refuses_large_tournaments(eng):
    from xtax_amd import DualParams
    from xtax_amd._native import XtxError
    P0 = workloads.lattice(8, 8, seed=3)[1]
    with pytest.raises(XtxError, match="target_samples"):
        eng.run_xdual(P0, None, EngineParams(replicas=2, max_epochs=1, target_samples=17),
                      DualParams())


def test_wide_batch_equals_alone(eng):
    mats = [workloads.planted(n, 100, 3 * n, seed=n)[0] for n in (12, 45, 133, 180, 201, 256)]
    seeds = [77 + q for q in range(len(mats))]
    p = EngineParams(replicas=4, t_init=60.0, t_min=10.0, exchange_interval=200, max_epochs=3)
    say("wide batch of 6 orders")
    res, st = eng.batch_solve(mats, p, seeds=seeds)
    for q, A0 in enumerate(mats):
        alone = eng.run_annealer(A0, None, EngineParams(
            replicas=4, t_init=60.0, t_min=10.0, exchange_interval=200, max_epochs=3,
            seed=seeds[q]))
        A, X, d = res[q]
        assert d["verify_status"] == 0 and d["solved"] == alone.stats["solved"], q
        assert np.array_equal(A, alone.A) and np.array_equal(X, alone.X), q


def test_wide_cli(tmp_path, orc):
    n = 140
    A0 = workloads.planted(n, 50, 2 * n, seed=3)[0]
    path = tmp_path / "A.csv"
    np.savetxt(path, A0, fmt="%d", delimiter=",")
    say("wide CLI n=140")
    out = subprocess.run([os.path.join(ROOT, "xtax_amd", "xtax"), "-A", str(path), "-t", "8",
                          "--max-epochs", "4", "--seed", "9"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    A = np.loadtxt(tmp_path / "best_A.csv", delimiter=",", dtype=np.int64)
    X = np.loadtxt(tmp_path / "best_X.csv", delimiter=",", dtype=np.int64)
    status, sign, _ = orc.verify(A0, X, A)
    assert status & ~2 == 0 and sign in (1, -1)
