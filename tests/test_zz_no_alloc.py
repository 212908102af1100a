"""The no-alloc plain solve (xtx_solve_plain_noalloc) against the plain solve
(xtx_solve_plain), through the C ABI.

Two calls in a row, the second on a smaller input (the scratch array keeps the
first call's contents), each bit-equal to the allocating entry point: the
returned A and X, every statistic but the timings, every replica's final
state and the move log. Sizes beyond the scratch array are a clear argument
error that leaves the context usable.
"""
import sys

import numpy as np
import pytest

from xtax_amd import EngineParams, workloads

pytestmark = pytest.mark.gpu

SAME = ("moves", "accepts", "greedy_calls", "infeasible", "sweeps", "exchanges_tried",
        "exchanges_accepted", "solved", "solve_replica", "solve_move", "best_score",
        "best_replica", "epochs", "stop_reason", "verify_status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _pair(eng, A0, p, R, n, log=False):
    ref = eng.run_annealer(A0, None, p)
    rA, rXt, rsc = eng.last_replicas(R, n)
    rlogs = [eng.movelog(r)[0] for r in range(R)] if log else None
    got = eng.run_annealer(A0, None, p, noalloc=True)
    gA, gXt, gsc = eng.last_replicas(R, n)
    glogs = [eng.movelog(r)[0] for r in range(R)] if log else None
    assert np.array_equal(got.A, ref.A) and np.array_equal(got.X, ref.X)
    for k in SAME:
        assert got.stats[k] == ref.stats[k], (k, got.stats[k], ref.stats[k])
    assert np.array_equal(gA, rA) and np.array_equal(gXt, rXt) and np.array_equal(gsc, rsc)
    if log:
        for r in range(R):
            assert np.array_equal(glogs[r], rlogs[r]), r
    return got


@pytest.mark.parametrize("look", [0, 1])
def test_noalloc_equals_plain_twice_second_smaller(eng, look):
    # first call: a planted order-40 form, 128 replicas (about 6.6 MB of state)
    A1, _, _ = workloads.planted(40, 100, 120, seed=21)
    p1 = EngineParams(replicas=128, seed=5, exchange_interval=300, max_epochs=3, t_init=18.0,
                      t_min=14.0, target_fraction=1.0, lookahead=look)
    say(f"noalloc: n=40 R=128 lookahead={look}")
    _pair(eng, A1, p1, 128, 40)
    # second call: smaller (order 12, 16 replicas) with the move log on, so the
    # scratch array still holds the larger call's state beyond what it writes
    A2, _, _ = workloads.planted(12, 50, 36, seed=22)
    p2 = EngineParams(replicas=16, seed=9, exchange_interval=200, max_epochs=4, t_init=30.0,
                      t_min=1.0, move_log=True, move_log_cap=4096, lookahead=look)
    say(f"noalloc: n=12 R=16 lookahead={look} (move log)")
    r2 = _pair(eng, A2, p2, 16, 12, log=True)
    assert r2.stats["verify_status"] == 0


def test_noalloc_rejects_what_does_not_fit(eng):
    from xtax_amd import _native as N
    A0, _, _ = workloads.planted(64, 100, 192, seed=3)
    p = EngineParams(replicas=2048, seed=1, max_epochs=1)  # about 270 MB of state
    say("noalloc: oversize n=64 R=2048 must be refused")
    with pytest.raises(N.XtxError) as ei:
        eng.run_annealer(A0, None, p, noalloc=True)
    assert ei.value.code == -1  # XTX_E_ARG
    assert "scratch" in str(ei.value)
    # the context still runs both entry points
    A2, _, _ = workloads.planted(10, 20, 30, seed=4)
    _pair(eng, A2, EngineParams(replicas=8, seed=2, exchange_interval=100, max_epochs=2,
                                t_init=20.0, t_min=1.0), 8, 10)
