"""Trajectory parity over randomly drawn engine configurations: for each of
24 configurations (order, matrix family, replicas, tempering or cooling,
greedy / target / sweep fractions, tournament size, stall threshold, ladder,
exchange interval, seed), drawn once from a fixed seed, a solve for three
epochs on the GPU ends in exactly the state the oracle's trajectory model
reaches: every replica's A and X^T (the solving replica's when one solves),
the scalars, and the returned best. The L1 objective on the first 16
configurations, xdual's objective on the last 8."""
import sys

import numpy as np
import pytest

from xtax_amd import DualParams, EngineParams, workloads

pytestmark = pytest.mark.gpu


def draw(k):
    rng = np.random.default_rng(1000 + k)
    n = int(rng.choice([3, 5, 8, 11, 16, 17, 24, 31, 33, 40, 48, 57, 64, 70, 90, 128]))
    fam = int(rng.integers(0, 3))
    if fam == 0:
        A0 = workloads.random_symmetric(n, -20, 20, seed=k)
    elif fam == 1:
        A0 = workloads.planted(n, 100, 3 * n, seed=k)[0]
    else:
        A0 = workloads.lattice(n, n + 2, seed=k)[1]
    kw = dict(replicas=int(rng.choice([1, 2, 3, 5, 8])), seed=int(rng.integers(0, 2 ** 40)),
              tempering=bool(rng.integers(0, 2)), exchange_interval=int(rng.integers(50, 400)),
              greedy_fraction=float(rng.choice([0.0, 0.3, 0.5, 1.0])),
              target_fraction=float(rng.choice([0.0, 0.5, 1.0])),
              target_samples=int(rng.choice([1, 3, 8, 16])),
              sweep_fraction=float(rng.choice([0.0, 0.2])),
              stuck_threshold=int(rng.choice([50, 300, 20000])),
              worker_diversity=float(rng.choice([0.0, 0.4])),
              t_init=float(rng.choice([0.0, 5.0, 50.0])), t_min=float(rng.choice([0.1, 2.0])),
              reseed_factor=float(rng.choice([1.0, 1.25])),
              adaptive_cooling=bool(rng.integers(0, 2)),
              moves_per_cool=int(rng.choice([20, 200])))
    return A0, kw


@pytest.fixture(scope="module")
def eng():
    from xtax_amd import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("k", range(16))
def test_l1_random_config(eng, k):
    from oracle import pyoracle as po
    A0, kw = draw(k)
    n = A0.shape[0]
    if np.count_nonzero(np.triu(A0, 1)) == 0:
        pytest.skip("diagonal start")
    look = [0, 1, 2, 4, 8, 16, 32][k % 7]  # proposals evaluated at once per replica (0: the rule)
    print("config", k, "n", n, "lookahead", look, kw, flush=True)
    sys.stdout.flush()
    res = eng.run_annealer(A0, None, EngineParams(max_epochs=3, lookahead=look, **kw),
                           allow_unverified=False)
    R = kw["replicas"]
    gA, gXt, gsc = eng.last_replicas(R, n)
    m = po.model_run(A0, np.eye(n, dtype=np.int64),
                     po.make_params(**{a: b for a, b in kw.items() if a in po.DEFAULTS}), 3)
    st = res.stats
    assert st["solved"] == m["solved"]
    if m["solved"]:
        r = m["solve_replica"]
        assert st["solve_replica"] == r and st["solve_move"] == m["solve_move"]
        assert np.array_equal(gA[r], m["A"][r]) and np.array_equal(gXt[r], m["Xt"][r])
    else:
        assert np.array_equal(gA, m["A"]) and np.array_equal(gXt, m["Xt"])
        assert np.array_equal(gsc, m["scal"])
        b = m["best_replica"]
        assert np.array_equal(res.A, m["best_A"][b]) and np.array_equal(res.X, m["best_Xt"][b].T)


@pytest.mark.parametrize("k", range(16, 24))
def test_dual_random_config(eng, k):
    from oracle import pyoracle as po
    A0, kw = draw(k)
    n = A0.shape[0]
    ok, Q0 = po.invert(A0)
    if not ok or np.count_nonzero(np.triu(A0, 1)) == 0:
        pytest.skip("singular or diagonal start")
    rng = np.random.default_rng(2000 + k)
    refresh = int(rng.choice([0, 150, 600])), int(rng.choice([0, 100, 400]))
    lam = float(rng.choice([0.5, 1.0, 2.0]))
    look = [0, 1, 2, 4, 8, 16][k % 6]  # lane groups per replica (0: the rule)
    print("dual config", k, "n", n, "lookahead", look, kw, refresh, lam, flush=True)
    res = eng.run_xdual(A0, None, EngineParams(max_epochs=3, lookahead=look, **kw),
                        DualParams(lambda_=lam, dual_refresh=refresh[0], dual_check=refresh[1],
                                   dual_tol=1e-9), allow_unverified=False)
    st = res.stats
    R = kw["replicas"]
    P, Q, Xt, dsc, sc = eng.dual_last_replicas(R, n)
    m = po.dual_model_run(A0, Q0, np.eye(n, dtype=np.int64),
                          po.make_params(**{a: b for a, b in kw.items() if a in po.DEFAULTS}),
                          po.DualParams(c=st["c"], dual_tol=1e-9, dual_refresh=refresh[0],
                                        dual_check=refresh[1]), 3)
    assert st["solved"] == m["solved"]
    reps = [m["solve_replica"]] if m["solved"] else range(R)
    for r in reps:
        assert np.array_equal(P[r], m["P"][r]) and np.array_equal(Xt[r], m["Xt"][r])
        assert np.array_equal(Q[r], m["Q"][r])
    if not m["solved"]:
        assert dsc.tolist() == m["dscal"].tolist() and np.array_equal(sc, m["scal"])
    assert np.array_equal(res.P, m["best_P"]) and np.array_equal(res.Q, m["best_Q"])
