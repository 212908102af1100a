"""Host-side parts of xtax around the annealer, through the C ABI (no device):
publish-time reordering (--rcm, --centroid), the unimodularity test of
--deflate and the strict-deflation pivot clearing, against the reference's own
outputs (tests/golden/g3_publish_deflate.npz, from src/main.cpp:244-350,
:407-411, :435-455 and src/mat_io.hpp:597-644) and against what the
mathematics fixes (permutation similarity, score and nonzero invariance,
det = +-1, X^T A X identities)."""
import ctypes as C
import os

import numpy as np
import pytest

from xtax_amd import _native as N

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "g3_publish_deflate.npz")


@pytest.fixture(scope="module")
def g3():
    return np.load(GOLD)


def P(a):
    return a.ctypes.data_as(N.I64P)


def I32(a):
    return a.ctypes.data_as(N.I32P)


def reorder(A, X, flags):
    A = np.ascontiguousarray(A, np.int64).copy()
    X = None if X is None else np.ascontiguousarray(X, np.int64).copy()
    perm = np.zeros(A.shape[0], np.int32)
    rc = N.lib().xtx_reorder(P(A), P(X) if X is not None else None, A.shape[0], flags, I32(perm))
    assert rc == 0
    return A, X, perm


def score(A):
    return 2 * int(np.abs(A).sum()) - int(np.abs(np.diag(A)).sum())


def offnz(A):
    return int(np.count_nonzero(np.triu(A, 1)))


def test_rcm_and_centroid_orders_match_reference(g3):
    for t in range(int(g3["reorder_cases"])):
        A = g3[f"r{t}_A"]
        _, _, p_rcm = reorder(A, None, 1)
        _, _, p_cen = reorder(A, None, 2)
        assert np.array_equal(p_rcm, g3[f"r{t}_rcm"]), t
        assert np.array_equal(p_cen, g3[f"r{t}_centroid"]), t


def test_reorder_for_publish_matches_reference(g3):
    """A <- A[p][p] and Xt rows permuted (X columns): the reference permutes
    the transposed transform, this ABI the transform itself."""
    for t in range(int(g3["reorder_cases"])):
        A, Xt = g3[f"r{t}_A"], g3[f"r{t}_Xt"]
        for flags in (1, 2, 3):
            A2, X2, perm = reorder(A, Xt.T, flags)
            assert np.array_equal(A2, g3[f"r{t}_A{flags}"]), (t, flags)
            assert np.array_equal(X2.T, g3[f"r{t}_Xt{flags}"]), (t, flags)
            # what a symmetric permutation fixes, independently of the reference
            assert sorted(perm.tolist()) == list(range(A.shape[0]))
            assert np.array_equal(A2, A[np.ix_(perm, perm)])
            assert score(A2) == score(A) and offnz(A2) == offnz(A)


def test_rcm_narrows_a_permuted_band():
    rng = np.random.default_rng(5)
    n = 30
    B = np.diag(np.full(n, 3)) + np.diag(np.ones(n - 1, np.int64), 1) + \
        np.diag(np.ones(n - 1, np.int64), -1)
    perm = rng.permutation(n)
    A = B[np.ix_(perm, perm)].astype(np.int64)
    A2, _, _ = reorder(A, None, 1)
    rows, cols = np.nonzero(A2)
    assert np.abs(rows - cols).max() == 1   # a path graph's RCM order is a path


def test_is_unimodular_matches_reference(g3):
    primes = (1000000007, 998244353)
    for u in range(int(g3["unimodular_cases"])):
        A = np.ascontiguousarray(g3[f"u{u}_A"])
        res = C.c_int32(-1)
        assert N.lib().xtx_is_unimodular(P(A), A.shape[0], C.byref(res)) == 0
        assert res.value == int(g3[f"u{u}_uni"]), u
        # the reference's rule from its own det residues
        d = g3[f"u{u}_det"]
        want = all(int(x) == 1 for x in d) or all(int(x) == p - 1 for x, p in zip(d, primes))
        assert res.value == int(want)


def test_is_unimodular_closed_forms():
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 12):
        U = np.eye(n, dtype=np.int64)
        for _ in range(4 * n):
            if n < 2:
                break
            i, j = rng.choice(n, 2, replace=False)
            U[:, j] += int(rng.choice([-1, 1])) * U[:, i]
        for d, want in ((np.ones(n, np.int64), 1), (np.r_[2, np.ones(n - 1, np.int64)], 0)):
            A = np.ascontiguousarray((U.T * d) @ U)
            res = C.c_int32(-1)
            assert N.lib().xtx_is_unimodular(P(A), n, C.byref(res)) == 0
            assert res.value == want


def test_deflate_pivot_matches_reference(g3):
    for d in range(int(g3["deflate_cases"])):
        M = np.ascontiguousarray(g3[f"d{d}_M"]).copy()
        Xt = np.ascontiguousarray(g3[f"d{d}_Xt"]).copy()
        M0, X0 = M.copy(), Xt.copy()
        i = int(g3[f"d{d}_i"])
        act = np.ascontiguousarray(g3[f"d{d}_active"], np.int32)
        cleared = C.c_int32(-1)
        rc = N.lib().xtx_deflate_pivot(P(M), P(Xt), M.shape[0], i, I32(act), len(act),
                                       C.byref(cleared))
        assert rc == 0
        assert cleared.value == int(g3[f"d{d}_ok"]), d
        assert np.array_equal(M, g3[f"d{d}_M2"]) and np.array_equal(Xt, g3[f"d{d}_Xt2"]), d
        if cleared.value:
            # row/column i cleared over the active set, the congruence kept:
            # M' = S^T M S with S = (X0^T)^{-1} X'^T unimodular
            others = [j for j in act if j != i]
            assert all(M[j, i] == 0 and M[i, j] == 0 for j in others)
            S = np.linalg.solve(X0.T.astype(float), Xt.T.astype(float))
            Si = np.rint(S).astype(np.int64)
            assert np.allclose(S, Si) and round(abs(np.linalg.det(Si))) == 1
            assert np.array_equal(Si.T @ M0 @ Si, M)
        else:
            assert np.array_equal(M, M0) and np.array_equal(Xt, X0)


def test_deflate_pivot_rejects_a_non_unit_pivot():
    M = np.array([[2, 1], [1, 3]], np.int64)
    Xt = np.eye(2, dtype=np.int64)
    act = np.array([0, 1], np.int32)
    cleared = C.c_int32(-1)
    assert N.lib().xtx_deflate_pivot(P(M), P(Xt), 2, 0, I32(act), 2, C.byref(cleared)) == -1
