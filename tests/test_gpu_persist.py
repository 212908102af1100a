"""The batch's persistent launch (k_lane_persist: waves take ready units of
any instance of its XCD, each instance's epoch end runs in the wave that
finishes its epoch) against the same batch run as captured graphs of epochs:
every instance's A, X, solve move, best score (and move count when unsolved),
and the batch's solve and verify totals, are equal."""
import numpy as np
import pytest

from xtax_amd import Engine, EngineParams, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def eng():
    with Engine(0) as e:
        yield e


CASES = {
    # the 1024-instance batch at its real shape (3 epochs)
    "batch1024": dict(count=1024, seed=7, reps=None,
                      kw=dict(replicas=32, t_init=18.0, t_min=14.0, target_fraction=1.0,
                              exchange_interval=2000, max_epochs=3)),
    # uneven replica counts (units with idle replica slots), tempering
    "uneven": dict(count=160, seed=5, reps=[25 + (q % 16) for q in range(160)],
                   kw=dict(replicas=32, t_init=18.0, t_min=14.0, target_fraction=1.0,
                           exchange_interval=500, max_epochs=6)),
    # cooling mode: stall sweeps (evaluated by the lane group), reseeds from
    # the instance-best snapshot the epoch end writes
    "cooling": dict(count=160, seed=9, reps=None,
                    kw=dict(replicas=32, tempering=False, t_init=40.0, stuck_threshold=300,
                            sweep_fraction=0.5, reseed_factor=1.0, exchange_interval=300,
                            max_epochs=6)),
    # more epochs than one launch runs (64 + 6)
    "relaunch": dict(count=128, seed=3, reps=None,
                     kw=dict(replicas=32, t_init=18.0, t_min=14.0, exchange_interval=100,
                             max_epochs=70)),
    # no epoch cap: the instances end when every replica reaches max_moves
    "uncapped": dict(count=128, seed=4, reps=None,
                     kw=dict(replicas=32, t_init=18.0, t_min=14.0, exchange_interval=500,
                             max_moves=3000)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_persistent_equals_epoch_graphs(eng, case):
    c = CASES[case]
    mats = workloads.batch_config(count=c["count"], seed=c["seed"])
    seeds = [7_000_001 + 13 * q for q in range(len(mats))]
    reps = c["reps"]
    say(f"persistent vs graphs: {case}, {len(mats)} instances")
    a, sa = eng.batch_solve(mats, EngineParams(**c["kw"]), seeds=seeds, replicas=reps)
    say(f"persistent: {sa['solved']} solved, {sa['kernel_ms']:.1f} ms")
    b, sb = eng.batch_solve(mats, EngineParams(persistent=False, **c["kw"]), seeds=seeds,
                            replicas=reps)
    say(f"graphs: {sb['solved']} solved, {sb['kernel_ms']:.1f} ms")
    assert sa["persistent"] == 1 and sb["persistent"] == 0
    assert sa["lane_instances"] == len(mats)
    for k in ("solved", "verified"):
        assert sa[k] == sb[k], k
    assert sa["solved"] == sa["verified"]
    for q in range(len(mats)):
        A, X, d = a[q]
        A2, X2, d2 = b[q]
        # a solved instance's move count depends on when its other replicas
        # see the solve, in either form
        keys = ("solved", "solve_move", "best_score", "verify_status") + (
            () if d["solved"] else ("moves",))
        for k in keys:
            assert d[k] == d2[k], (case, q, k)
        assert np.array_equal(A, A2) and np.array_equal(X, X2), (case, q)


def test_persistent_falls_back_when_a_shape_is_set(eng):
    """An order set to a lane shape other than the rule's runs the batch as
    graphs of epochs, with the same results."""
    mats = workloads.batch_config(count=128, seed=2)
    seeds = [11 + q for q in range(len(mats))]
    kw = dict(replicas=32, t_init=18.0, t_min=14.0, exchange_interval=400, max_epochs=3)
    a, sa = eng.batch_solve(mats, EngineParams(**kw), seeds=seeds)
    eng.launch_set(40, 1, 40, 1, 0, 16)
    try:
        b, sb = eng.batch_solve(mats, EngineParams(**kw), seeds=seeds)
    finally:
        eng.launch_reset(40)
    assert sa["persistent"] == 1 and sb["persistent"] == 0
    for q in range(len(mats)):
        assert np.array_equal(a[q][0], b[q][0]) and np.array_equal(a[q][1], b[q][1]), q


def test_persistent_batch_stops_at_its_deadline(eng):
    """A deadline ends a persistent launch: the waves leave at their next poll
    (a stop sets every die's live count far below zero, so waves waiting for
    work leave too), the call returns with every instance retired and
    verified, and the next batch on the context runs in full."""
    import time
    mats = workloads.batch_config(count=512, seed=13)
    seeds = [5_000_011 + q for q in range(len(mats))]
    kw = dict(replicas=32, t_init=18.0, t_min=14.0, exchange_interval=2000)
    t = time.perf_counter()
    out, st = eng.batch_solve(mats, EngineParams(max_seconds=0.3, **kw), seeds=seeds)
    wall = time.perf_counter() - t
    say(f"deadline batch: {st['solved']} solved, {st['epochs']} epochs, {wall:.2f} s")
    assert st["persistent"] == 1 and wall < 5.0
    assert st["solved"] == st["verified"]
    assert all(d["verify_status"] == 0 for _, _, d in out)
    # the stop word is clear again: a capped batch runs all its epochs
    out2, st2 = eng.batch_solve(mats[:128], EngineParams(max_epochs=2, **kw), seeds=seeds[:128])
    assert st2["persistent"] == 1
    assert all(d["moves"] == 2 * 2000 * 32 for _, _, d in out2 if not d["solved"])
