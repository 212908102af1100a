"""The benchmark configurations at their real shapes on the GPU, against the
oracle's trajectory model and against the plain form.

* Plain matrix n = 10 (random symmetric, entries U[-20, 20]), 8 tempering
  replicas on the ladder [0.5, 50], greedy 0.5: four seeds, every replica's
  state and move log equal to the model's.
* Lattice 24 x 24 (basis U[-9, 9]), Gram in int64, 256 replicas, greedy 0.5,
  sweep 0.05: two epochs equal to the model.
* Replica exchange past one staging chunk of the epoch-end kernel (it stages
  512 rungs at a time): 513, 1100 and 4096 replicas at order 8 with short
  epochs, every replica's state, rung and the exchange count equal to the
  model's.
* Large planted n = 80 (D = +-U{1..1000}, 4n shears), 4096 replicas, greedy
  0.7: two epochs with a short exchange interval equal to the model, with
  each replica's X^T in HBM (the rule) and in LDS (xtx_launch x_in_lds);
  the LDS form also on cooling runs with sweeps and reseeds, at orders 72 to
  112 (the largest whose A and X^T fit one workgroup's LDS), and refused at
  order 120.
* The full batch: 1024 planted instances, n ~ U{8..48}, 32 replicas each, for
  three epochs: every solved result verified exactly, 32 sampled instances
  equal to their plain runs, four of them equal to the model.
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


def _both(**kw):
    from oracle import pyoracle
    ep = EngineParams(**kw)
    op = pyoracle.make_params(**{k: v for k, v in kw.items() if k in pyoracle.DEFAULTS})
    return ep, op


def _trajectory(eng, orc, A0, kw, epochs, log_cap=0):
    """Plain solve vs the model for `epochs` epochs: every replica's state,
    scalars, the returned result and (with a log) every accepted move."""
    n = A0.shape[0]
    ep, op = _both(max_epochs=epochs, move_log=log_cap > 0, move_log_cap=log_cap, **kw)
    res = eng.run_annealer(A0, None, ep)
    model = orc.model_run(A0, np.eye(n, dtype=np.int64), op, epochs, log_cap=log_cap)
    st = res.stats
    R = ep.replicas
    gA, gXt, gsc = eng.last_replicas(R, n)
    assert st["solved"] == model["solved"] and st["verify_status"] == 0
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
        assert np.array_equal(gsc, model["scal"]), "replica scalars differ"
        b = model["best_replica"]
        assert st["best_replica"] == b and st["best_score"] == model["best_score"]
        assert np.array_equal(res.A, model["best_A"][b])
        assert np.array_equal(res.X, model["best_Xt"][b].T)
        assert st["exchanges_accepted"] == model["exchanges"]
        reps = range(R)
    if log_cap:
        for r in reps:
            glog, trunc = eng.movelog(r)
            assert not trunc and np.array_equal(glog, model["logs"][r]), r
    return res, model


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_plain_n10_random_symmetric(eng, orc, seed):
    A0 = workloads.random_symmetric(10, -20, 20, seed=seed)
    say(f"config plain n=10 seed={seed}")
    _trajectory(eng, orc, A0, dict(replicas=8, seed=seed, t_init=50.0, t_min=0.5,
                                   greedy_fraction=0.5), epochs=3, log_cap=20000)


def test_lattice_24(eng, orc):
    _, A0 = workloads.lattice(24, 24, -9, 9, seed=1)
    say("config lattice 24x24, 256 replicas")
    _trajectory(eng, orc, A0, dict(replicas=256, seed=3, greedy_fraction=0.5,
                                   sweep_fraction=0.05, exchange_interval=1000), epochs=2)


@pytest.mark.parametrize("R", [513, 1100, 4096])
def test_exchange_past_one_staging_chunk(eng, orc, R):
    # a random symmetric form (not diagonalized here), so every replica's
    # state and rung is compared
    A0 = workloads.random_symmetric(8, -20, 20, seed=R)
    say(f"config exchange R={R}")
    res, model = _trajectory(eng, orc, A0, dict(replicas=R, seed=R, t_init=60.0, t_min=2.0,
                                                exchange_interval=40), epochs=4)
    assert not model["solved"]
    assert model["exchanges"] > R  # swaps accepted on both sides of every chunk boundary


@pytest.mark.parametrize("xl", [0, 1])
def test_large_n80_4096_replicas(eng, orc, xl):
    A0, _, _ = workloads.planted(80, 1000, 320, seed=1)
    say(f"config large n=80, 4096 replicas, 2 epochs of 100 moves, X^T in LDS={xl}")
    if xl:
        eng.launch_set(80, 64, 2, 1, 0, 4, x_in_lds=1)
        cfg = eng.launch_get(80)
        assert (cfg["x_in_lds"], cfg["lookahead"], cfg["tuned"]) == (1, 4, 1)
        assert cfg["lds_bytes"] >= 80 * 80 * 8 + 80 * 81 * 4
    try:
        _trajectory(eng, orc, A0, dict(replicas=4096, seed=5, greedy_fraction=0.7,
                                       t_init=2000.0, t_min=50.0, exchange_interval=100),
                    epochs=2)
    finally:
        eng.launch_reset(80)


@pytest.mark.parametrize("n", [72, 80, 97, 112])
def test_x_in_lds_sweeps_reseeds(eng, orc, n):
    """X^T in LDS through the paths that read and write it outside a commit:
    stall sweeps, reseeds from the instance best, best flushes, move logs."""
    A0, _, _ = workloads.planted(n, 300, 2 * n, seed=n)
    order = -(-n // 8) * 8
    eng.launch_set(order, 64, 2, 1, 0, 4, x_in_lds=1)
    say(f"X^T in LDS: n={n} cooling, sweeps, reseeds")
    try:
        _trajectory(eng, orc, A0, dict(replicas=5, seed=n, tempering=False, t_init=40.0,
                                       exchange_interval=300, stuck_threshold=60,
                                       sweep_fraction=0.5, reseed_factor=1.0,
                                       greedy_fraction=0.7),
                    epochs=3, log_cap=4000)
    finally:
        eng.launch_reset(order)


def test_x_in_lds_refused_where_it_does_not_fit(eng):
    from xtax_amd._native import XtxError
    with pytest.raises(XtxError, match="LDS"):
        eng.launch_set(120, 64, 2, 1, 0, 4, x_in_lds=1)


def test_full_batch_1024(eng, orc):
    mats = workloads.batch_config(count=1024, seed=7)
    seeds = [1_000_003 + q for q in range(len(mats))]
    kw = dict(replicas=32, t_init=18.0, t_min=14.0, greedy_fraction=0.5, target_fraction=1.0,
              exchange_interval=2000, max_epochs=3)
    say("config batch 1024 x 32 replicas, 3 epochs")
    out, bst = eng.batch_solve(mats, EngineParams(**kw), seeds=seeds)
    say(f"batch: {bst['solved']} solved, {bst['verified']} verified, {bst['epochs']} epochs")
    assert bst["instances"] == 1024 and bst["solved"] == bst["verified"]
    for q, (A, X, d) in enumerate(out):
        if d["solved"]:
            assert d["verify_status"] == 0, q
            assert np.count_nonzero(A - np.diag(np.diag(A))) == 0
    rng = np.random.default_rng(3)
    sample = sorted(rng.choice(1024, 32, replace=False).tolist())
    for k, q in enumerate(sample):
        r = eng.run_annealer(mats[q], None, EngineParams(**{**kw, "seed": seeds[q]}))
        A, X, d = out[q]
        assert d["solved"] == r.stats["solved"], q
        if d["solved"]:
            assert d["solve_move"] == r.stats["solve_move"], q
        assert np.array_equal(A, r.A) and np.array_equal(X, r.X), q
        if k < 4:
            n = mats[q].shape[0]
            _, op = _both(**{**kw, "seed": seeds[q]})
            m = orc.model_run(mats[q], np.eye(n, dtype=np.int64), op, 3)
            assert d["solved"] == m["solved"]
            b = m["solve_replica"] if m["solved"] else m["best_replica"]
            want = m["A"][b] if m["solved"] else m["best_A"][b]
            assert np.array_equal(A, want), q
