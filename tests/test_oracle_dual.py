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
    for n in (1, 2, 3, 7, 64, 100, 128):
        x = rng.standard_normal(n)
        buf = np.concatenate([x, np.zeros(128 - n)])
        while len(buf) > 1:
            buf = buf[0::2] + buf[1::2]
        assert po.pairsum(x) == buf[0]
    assert po.pairsum(np.array([2.0 ** 60, 1.0, -2.0 ** 60, 1.0])) == 0.0  # (a+1)+(-a+1) rounds


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
