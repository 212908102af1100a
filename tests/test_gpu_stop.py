"""The stop path (reference: stop_signal.hpp:40-89, congruence_anneal.hpp:
502-519, main.cpp:725-731).

1. xtx_request_stop from a second host thread ends a running plain solve at
   its next poll with stop_reason 2; the best state so far comes back
   verified; the request stays until xtx_clear_stop, after which solves run.
2. A solve's own max_seconds deadline never outlives the solve: whenever it
   fires (before, while or after the loop ends), the next solve on the
   context runs to its own cap.
3. The command line: the first SIGINT stops a run gracefully, and xtax exits
   0 after writing best_X.csv and best_A.csv, which pass the exact verifier.
"""
import os
import signal
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

from xtax_amd import EngineParams, workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XTAX = os.path.join(ROOT, "xtax_amd", "xtax")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _hard():
    # a random symmetric form of order 10: these settings do not diagonalize it
    return workloads.random_symmetric(10, seed=3)


def test_request_stop_from_another_thread(eng, orc):
    A0 = _hard()
    p = EngineParams(replicas=64, seed=4, exchange_interval=2000)  # no cap of any kind
    timer = threading.Timer(1.0, eng.request_stop)
    say("stop: solve with no cap, stop requested after 1 s")
    timer.start()
    try:
        r = eng.run_annealer(A0, None, p)
    finally:
        timer.cancel()
    say(f"stop: returned after {r.stats['seconds']:.2f} s, {r.stats['epochs']} epochs")
    assert r.stats["stop_reason"] == 2 and r.stats["solved"] == 0
    assert 0.9 < r.stats["seconds"] < 30.0
    assert r.stats["verify_status"] == 0
    status, sign, _ = orc.verify(A0, r.X, r.A)
    assert status & ~2 == 0 and sign in (1, -1)
    # the request holds until cleared (the reference's g_stop is sticky too)
    r2 = eng.run_annealer(A0, None, EngineParams(replicas=8, seed=5, max_epochs=50))
    assert r2.stats["stop_reason"] == 2 and r2.stats["epochs"] == 1
    eng.clear_stop()
    r3 = eng.run_annealer(A0, None, EngineParams(replicas=8, seed=5, exchange_interval=500,
                                                 max_epochs=3))
    assert r3.stats["stop_reason"] == 3 and r3.stats["epochs"] == 3


def test_deadline_never_outlives_its_solve(eng):
    A0 = _hard()
    base = EngineParams(replicas=256, seed=6, exchange_interval=2000, max_epochs=3)
    t = eng.run_annealer(A0, None, base).stats["seconds"]
    say(f"deadline: 3 epochs take {t:.3f} s")
    for frac in (0.25, 0.5, 0.8, 0.95, 1.0, 1.05, 1.2):
        p = EngineParams(**{**base.__dict__, "max_seconds": frac * t})
        r = eng.run_annealer(A0, None, p)
        assert r.stats["stop_reason"] in (1, 3), (frac, r.stats["stop_reason"])
        time.sleep(0.05)
        nxt = eng.run_annealer(A0, None, EngineParams(replicas=8, seed=7, exchange_interval=300,
                                                      max_epochs=4))
        assert nxt.stats["stop_reason"] == 3 and nxt.stats["epochs"] == 4, (frac, nxt.stats)
    # the batch form's watchdog outlives its loop (retirement, verification)
    mats = [workloads.planted(n, 100, 3 * n, seed=n)[0] for n in (12, 20, 28, 36)] * 8
    bp = EngineParams(replicas=32, t_init=18.0, t_min=14.0, exchange_interval=2000, max_epochs=2)
    tb = eng.batch_solve(mats, bp, seeds=list(range(len(mats))))[1]["seconds"]
    for frac in (0.5, 0.9, 1.0, 1.1):
        eng.batch_solve(mats, EngineParams(**{**bp.__dict__, "max_seconds": frac * tb}),
                        seeds=list(range(len(mats))))
        nxt = eng.run_annealer(A0, None, EngineParams(replicas=8, seed=7, exchange_interval=300,
                                                      max_epochs=4))
        assert nxt.stats["stop_reason"] == 3 and nxt.stats["epochs"] == 4, (frac, nxt.stats)


def _read_csv(path):
    return np.array([[int(x) for x in r.split(",")] for r in open(path).read().split("\n")
                     if r.strip()], np.int64)


def test_cli_sigint_writes_verified_best(tmp_path, orc):
    A0 = _hard()
    with open(tmp_path / "a.csv", "w") as f:
        for row in A0:
            f.write(",".join(str(int(v)) for v in row) + "\n")
    say("stop: xtax -A a.csv -t 64, SIGINT after the first report")
    proc = subprocess.Popen([XTAX, "-A", "a.csv", "-t", "64", "--seed", "9"], cwd=tmp_path,
                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        line = proc.stdout.readline()  # "[t=0s] start score=..." once it is running
        assert "start" in line, line
        time.sleep(1.5)
        proc.send_signal(signal.SIGINT)
        out, err = proc.communicate(timeout=60)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
    assert proc.returncode == 0, (proc.returncode, err[-2000:])
    assert "interrupt received" in err
    assert "[interrupted] writing best_X.csv and best_A.csv before exit" in out
    A = _read_csv(tmp_path / "best_A.csv")
    X = _read_csv(tmp_path / "best_X.csv")
    status, sign, _ = orc.verify(A0, X, A)
    assert status & ~2 == 0 and sign in (1, -1)
    assert orc.score(A) <= orc.score(A0)
