"""The served form (xtx_serve_*) under admission pressure, and forced
retirement.

1. A stream of planted serving requests (n = round(exp(U[ln 6, ln 64])),
   D = +-U{1..200}, 3n shears, 64 replicas each) arriving a few per epoch
   into a server whose LDS budget, resident-request cap and HBM budget are
   far below the stream: requests queue, and slots and arena blocks are
   reused many times over. Every result equals the same request solved alone
   (xtx_solve_plain with the same seed and parameters), bit for bit.
2. xtx_serve_retire on a request admitted at the end of a step (its record is
   not yet on the device, and its slot held another request a moment
   before): the result carries its own tag and its own start state, and the
   other requests still equal their alone runs.
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


# the serving workload's ladder and targeting, short epochs and an epoch cap
SERVE = dict(replicas=64, t_init=50.0, t_min=36.0, target_fraction=1.0, exchange_interval=250,
             max_epochs=6)


def _alone(eng, A0, seed, ep):
    r = eng.run_annealer(A0, None, EngineParams(**{**ep, "seed": seed}))
    return r.A, r.X, r.stats


def _drain(eng, got, limit=100000):
    for _ in range(limit):
        r, q = eng.serve_step()
        for A, X, d in eng.serve_poll():
            got[d["tag"]] = (A, X, d)
        if r == 0 and q == 0:
            return
    raise AssertionError("server did not drain")


def test_serve_under_pressure_equals_alone(eng):
    count = 480
    reqs = workloads.serve_config(count=count, dmax=200, seed=31)
    seeds = [7_000_003 * (q + 1) for q in range(count)]
    rng = np.random.default_rng(5)
    # 10 MB of LDS (a request holds 0.8 to 1.5 MB: 64 replicas at lookahead
    # 16 or 8), at most 8 resident and 160 MB of HBM: the stream queues
    eng.serve_open(EngineParams(**SERVE), lds_budget=10 << 20, hbm_budget=160 << 20,
                   max_resident=8)
    got = {}
    peak_res = peak_q = 0
    nxt = 0
    try:
        while nxt < count:
            for _ in range(int(rng.poisson(3.0))):
                if nxt < count:
                    eng.serve_submit(reqs[nxt][1], nxt, seeds[nxt])
                    nxt += 1
            r, q = eng.serve_step()
            peak_res, peak_q = max(peak_res, r), max(peak_q, q)
            for A, X, d in eng.serve_poll():
                got[d["tag"]] = (A, X, d)
        _drain(eng, got)
    finally:
        eng.serve_close()
    say(f"serve pressure: {count} requests, peak resident {peak_res}, peak queued {peak_q}")
    assert len(got) == count
    assert peak_res <= 8 and peak_q > 0
    solved = 0
    for q in range(count):
        A, X, d = got[q]
        aA, aX, ast = _alone(eng, reqs[q][1], seeds[q], SERVE)
        assert d["verify_status"] == 0, (q, d)
        assert d["solved"] == ast["solved"], q
        if ast["solved"]:
            assert d["solve_move"] == ast["solve_move"], q
        assert np.array_equal(A, aA) and np.array_equal(X, aX), q
        solved += d["solved"]
    say(f"serve pressure: {solved} of {count} solved, every result equal to its alone run")


def test_serve_retire_request_admitted_at_step_end(eng):
    # four hard requests (random symmetric forms, which these settings do not
    # diagonalize) hold the four slots until their epoch cap, then retire in
    # the same step; the queued requests take their slots in that step's
    # trailing admission, and one of them is cancelled at once
    ep = dict(SERVE, max_epochs=3, replicas=32)
    hard = [workloads.random_symmetric(10, seed=40 + q) for q in range(4)]
    easy = [workloads.planted(n, 50, 3 * n, seed=60 + n)[0] for n in (7, 9, 12, 16)]
    eng.serve_open(EngineParams(**ep), max_resident=4)
    got = {}
    try:
        for q, A0 in enumerate(hard):
            eng.serve_submit(A0, q, 100 + q)
        r, qd = eng.serve_step()
        assert (r, qd) == (4, 0)
        for q, A0 in enumerate(easy):
            eng.serve_submit(A0, 10 + q, 200 + q)
        r, qd = eng.serve_step()
        assert (r, qd) == (4, 4)
        r, qd = eng.serve_step()
        # the hard requests reached their cap in this step: their slots went
        # to the four queued requests, admitted after the retirements
        assert (r, qd) == (4, 0), (r, qd)
        for A, X, d in eng.serve_poll():
            got[d["tag"]] = (A, X, d)
        assert sorted(got) == [0, 1, 2, 3]
        eng.serve_retire(12)
        out = eng.serve_poll()
        assert len(out) == 1
        A, X, d = out[0]
        assert d["tag"] == 12 and d["solved"] == 0 and d["moves"] == 0, d
        assert d["verify_status"] == 0
        assert np.array_equal(A, easy[2]) and np.array_equal(X, np.eye(12, dtype=np.int64))
        _drain(eng, got)
    finally:
        eng.serve_close()
    for q in range(4):
        A, X, d = got[q]
        aA, aX, ast = _alone(eng, hard[q], 100 + q, ep)
        assert d["solved"] == ast["solved"] and np.array_equal(A, aA) and np.array_equal(X, aX)
    for q in (0, 1, 3):
        A, X, d = got[10 + q]
        aA, aX, ast = _alone(eng, easy[q], 200 + q, ep)
        assert d["solved"] == ast["solved"] and d["verify_status"] == 0
        assert np.array_equal(A, aA) and np.array_equal(X, aX), q
