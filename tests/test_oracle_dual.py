"""The oracle's xdual primitives (oracle/oracle_dual.cpp) against the
reference's own (tests/golden/g4_xdual.npz, from src/xdual_main.cpp:67-364 and
src/mat_io.hpp:297-337) and against exact arithmetic.

The reference accumulates the score deltas in long double; the engine's
specification (which the oracle restates) sums pairwise in double with the
primal squares formed exactly. So: the inverse and the proposed rows are
bit-exact; the deltas agree to a relative 1e-12 of their term magnitudes; the
best shear agrees except where its two candidates tie to within rounding,
which is checked with exact rational arithmetic."""
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "g4_xdual.npz")


@pytest.fixture(scope="module")
def g4():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def po():
    from oracle import pyoracle
    return pyoracle


def cases(g4):
    return range(int(g4["cases"]))


def test_pairwise_sum_closed_form(po):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 64, 65, 100, 128):
        x = rng.standard_normal(n)
        pad = np.concatenate([x, np.zeros(128 - n)])
        buf = pad[:64] + pad[64:]          # column k + 64 folds onto column k
        while len(buf) > 1:
            buf = buf[0::2] + buf[1::2]    # then adjacent pairs
        assert po.pairsum(x) == buf[0]
    assert po.pairsum(np.array([2.0 ** 60, 1.0, -2.0 ** 60, 1.0])) == 0.0  # (a+1)+(-a+1) rounds
    # the fold: x[0] + x[64] first, so 2^60 absorbs the 1 before -2^60 cancels it
    x = np.zeros(65)
    x[0], x[1], x[64] = 2.0 ** 60, -2.0 ** 60, 1.0
    assert po.pairsum(x) == 0.0


def test_invert_is_the_references(g4, po):
    for t in cases(g4):
        ok, Q = po.invert(g4[f"c{t}_P"])
        assert ok == bool(g4[f"c{t}_ok"])
        if ok:
            assert np.array_equal(Q, g4[f"c{t}_Q"]), t   # bit for bit
            P = g4[f"c{t}_P"].astype(float)
            assert np.allclose(P @ Q, np.eye(len(P)), atol=1e-6 * max(1.0, np.abs(P).max()))


def test_norms(g4, po):
    for t in cases(g4):
        P, Q = g4[f"c{t}_P"], g4[f"c{t}_Q"]
        exact = sum(int(v) ** 2 for r, row in enumerate(P.tolist()) for k, v in enumerate(row)
                    if r != k)
        assert po.frob2_off_p(P) == pytest.approx(float(exact), rel=1e-14)
        assert po.frob2_off_p(P) == pytest.approx(float(g4[f"c{t}_fp"]), rel=1e-14)
        assert po.frob2_off_q(Q) == pytest.approx(float(g4[f"c{t}_fq"]), rel=1e-13)


def test_eval_matches_reference(g4, po):
    for t in cases(g4):
        P, Q = g4[f"c{t}_P"], g4[f"c{t}_Q"]
        n = P.shape[0]
        for q, (i, j, s) in enumerate(g4[f"c{t}_q"].tolist()):
            ok, dp, dq, dnz, prow, qrow = po.dual_eval(P, Q, i, j, s)
            rok, rdnz = g4[f"c{t}_ev"][q].tolist()
            assert ok == bool(rok), (t, q)
            if not ok:
                continue
            assert dnz == rdnz
            # the proposed rows: P exactly (integers), Q bit for bit (same ops)
            want = [int(P[j, k]) + s * int(P[i, k]) for k in range(n)]
            want[j] = int(P[j, j]) + 2 * s * int(P[j, i]) + s * s * int(P[i, i])
            assert prow.tolist() == want
            assert np.array_equal(qrow, g4[f"c{t}_rows_q"][q]), (t, q)
            # primal delta: each term exact then rounded, the sum pairwise, so
            # within n * 2^-53 of the terms' magnitudes of the exact integer
            exact = 2 * sum(want[k] ** 2 - int(P[j, k]) ** 2 for k in range(n) if k != j)
            scale_p = 2 * sum(want[k] ** 2 + int(P[j, k]) ** 2 for k in range(n) if k != j)
            assert abs(Fraction(dp) - exact) <= Fraction(8 * n, 2 ** 53) * scale_p + 1
            rdp, rdq = g4[f"c{t}_evd"][q].tolist()
            assert abs(dp - rdp) <= 1e-12 * scale_p + 1.0
            scale_q = 2 * float(np.sum(qrow ** 2) + np.sum(Q[i] ** 2))
            assert abs(dq - rdq) <= 1e-12 * scale_q, (t, q, dq, rdq)


def _f_exact(P, Q, c, i, j, s):
    """f(s) = 2 (quad s^2 + 2 lin s) in exact rationals from the double Q."""
    n = P.shape[0]
    a2 = sum(Fraction(int(P[i, k])) ** 2 for k in range(n) if k != j)
    b2 = sum(Fraction(int(P[j, k]) * int(P[i, k])) for k in range(n) if k != j)
    c2 = sum(Fraction(float(Q[j, k])) ** 2 for k in range(n) if k != i)
    d2 = sum(Fraction(float(Q[i, k])) * Fraction(float(Q[j, k])) for k in range(n) if k != i)
    cf = Fraction(c)
    quad, lin = a2 + cf * c2, b2 - cf * d2
    return 2 * (quad * s * s + 2 * lin * s), quad


def test_best_shear_matches_reference(g4, po):
    total = mismatch = 0
    for t in cases(g4):
        P, Q, c = g4[f"c{t}_P"], g4[f"c{t}_Q"], float(g4[f"c{t}_c"])
        n = P.shape[0]
        best = g4[f"c{t}_best"]
        for i in range(n):
            for j in range(n):
                if i == j:
                    continue
                total += 1
                got = po.dual_best_shear(P, Q, c, i, j)
                if got == int(best[i, j]):
                    continue
                mismatch += 1
                # only a rounding-level tie between the two choices may differ
                fg, quad = _f_exact(P, Q, c, i, j, got)
                fr, _ = _f_exact(P, Q, c, i, j, int(best[i, j]))
                assert abs(fg - fr) <= Fraction(1, 10 ** 9) * (abs(quad) * 64 * 64 + 1), (t, i, j)
    assert mismatch <= total // 1000, (mismatch, total)


def test_residual_of_a_true_inverse(po):
    rng = np.random.default_rng(8)
    L = rng.integers(-5, 6, (9, 9))
    P = (L @ L.T).astype(np.int64)
    ok, Q = po.invert(P)
    assert ok
    assert all(po.dual_residual(P, Q, k) < 1e-9 for k in range(9))
    Q2 = Q.copy()
    Q2[3, 3] += 1e-3
    assert po.dual_residual(P, Q2, 3) > 1e-6


def _brute_best(P, Q, c, i, j):
    """argmin over s in [-64, 64] of the exact (rational) change f(s) < 0,
    the smaller s on a tie, else 0: the integer minimizer the best shear's
    floor/ceil rule must find for a convex quadratic."""
    best_s, best_f = 0, Fraction(0)
    for s in range(-64, 65):
        if s == 0:
            continue
        f, _ = _f_exact(P, Q, c, i, j, s)
        if f < best_f:
            best_s, best_f = s, f
    return best_s, best_f


def test_best_shear_is_the_integer_minimizer(g4, po):
    """Independently of the reference: at c = 0 the change is the exact
    integer change of |offdiag P|^2; at c > 0 the rational change of the
    double state. A differing choice must tie to rounding."""
    rng = np.random.default_rng(21)
    for t in list(cases(g4))[:18]:
        P, Q = g4[f"c{t}_P"], g4[f"c{t}_Q"]
        n = P.shape[0]
        for _ in range(12):
            i, j = (int(v) for v in rng.choice(n, 2, replace=False))
            for c in (0.0, float(g4[f"c{t}_c"])):
                got = po.dual_best_shear(P, Q, c, i, j)
                want, fw = _brute_best(P, Q, c, i, j)
                if c == 0.0:
                    # exact integer objective change for the chosen shear
                    ok, dp, _, _, prow, _ = po.dual_eval(P, Q, i, j, want) if want else (1, 0, 0, 0, 0, 0)
                    if want:
                        exact = 2 * sum(int(prow[k]) ** 2 - int(P[j, k]) ** 2 for k in range(n) if k != j)
                        assert Fraction(exact) == fw
                if got != want:
                    fg, quad = _f_exact(P, Q, c, i, j, got) if got else (Fraction(0), 1)
                    assert abs(fg - fw) <= Fraction(1, 10 ** 9) * (abs(quad) * 64 * 64 + 1), (t, i, j, c)


def _model_case(po, n, seed, tempering, R=4, epochs=4, E=300, refresh=700, check=250, tol=1e-9):
    rng = np.random.default_rng(seed)
    B = np.eye(n, dtype=np.int64)
    for _ in range(3 * n):
        a, b = rng.choice(n, 2, replace=False)
        B[a] += int(rng.integers(-2, 3)) * B[b]
    P0 = (B @ np.diag(rng.integers(1, 4, n)) @ B.T).astype(np.int64)
    ok, Q0 = po.invert(P0)
    assert ok
    p = po.Params(**dict(po.DEFAULTS, replicas=R, seed=seed, tempering=tempering,
                         exchange_interval=E, stuck_threshold=150, sweep_fraction=0.5,
                         greedy_fraction=0.6))
    fp, fq = po.frob2_off_p(P0), po.frob2_off_q(Q0)
    dp = po.DualParams(c=fp / fq, dual_tol=tol, dual_refresh=refresh, dual_check=check)
    out = po.dual_model_run(P0, Q0, np.eye(n, dtype=np.int64), p, dp, epochs, log_cap=4096)
    return P0, Q0, dp, out


@pytest.mark.parametrize("tempering", [1, 0])
def test_dual_model_invariants(po, tempering):
    """What the mathematics fixes about every replica of the model: P is the
    exact congruence X^T P0 X by a unimodular X, the running score tracks the
    recomputed one, Q stays an inverse of P, and (tempering: no reseeds) the
    accepted-move log replays P exactly through the move primitives."""
    n = 9
    P0, Q0, dp, out = _model_case(po, n, 5, tempering)
    assert out["refreshes"] > 0 and out["checks"] > 0
    for r in range(len(out["P"])):
        P, Q, Xt = out["P"][r], out["Q"][r], out["Xt"][r]
        Xo = Xt.T.astype(object)
        assert np.array_equal((Xo.T @ P0.astype(object) @ Xo).astype(np.int64), P)
        assert round(abs(np.linalg.det(Xt.astype(float)))) == 1
        assert int(out["scal"][r][0]) == int(np.count_nonzero(np.triu(P, 1)))
        # the dual drifts between re-inversions (condition numbers here reach 1e8)
        assert np.abs(P.astype(float) @ Q - np.eye(n)).max() < 1e-4
        score = out["dscal"][r][0]
        assert score == pytest.approx(po.dual_score(P, Q, dp.c), rel=1e-9)
        if tempering:
            Pr, Qr = P0.copy(), Q0.copy()
            for e in out["logs"][r]:
                i, j, s = int(e["i"]), int(e["j"]), int(e["s"])
                ok, _, _, _, prow, qrow = po.dual_eval(Pr, Qr, i, j, s)
                assert ok
                Pr[j, :] = prow
                Pr[:, j] = prow
                Qr[i, :] = qrow
                Qr[:, i] = qrow
            assert np.array_equal(Pr, P)


def test_dual_model_is_deterministic_and_solves(po):
    a = _model_case(po, 7, 3, 1, epochs=30)[3]
    b = _model_case(po, 7, 3, 1, epochs=30)[3]
    assert np.array_equal(a["P"], b["P"]) and np.array_equal(a["Q"], b["Q"])
    assert a["solved"] == 1
    bp = a["best_P"]
    assert np.count_nonzero(bp - np.diag(np.diag(bp))) == 0
