"""Seeded synthetic inputs for the annealer's benchmark configurations.

Draw order is fixed here (numpy PCG64 seeded per instance), so every input is
reproducible from (config, seed):

* random_symmetric(n, lo, hi, seed): upper triangle i <= j iid uniform in
  [lo, hi], row-major, mirrored. Config "plain matrix n=10": n=10, [-20, 20].
* lattice(m, d, lo, hi, seed): basis rows iid uniform in [lo, hi], redrawn
  while singular; Gram = L L^T (int64). Config "lattice 24x24": [-9, 9].
* planted(n, dmax, shears, seed): D = diag(+-U{1..dmax}) with a fair-coin sign,
  U = product of `shears` random unimodular shears (uniform ordered pair i != j,
  s = +-1: column j += s * column i), A = U^T D U, redrawn if any |A| > 2^48.
  A diagonal form exists (X = U^{-1}). Configs: batch (n ~ U{8..48}, dmax 100,
  3n shears), large (n = 80, dmax 1000, 4n shears), serve (n =
  round(exp(U[ln 6, ln 64])), dmax 200, 3n shears).
"""
import math

import numpy as np

LIMIT = 1 << 48


def random_symmetric(n, lo=-20, hi=20, seed=1):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n), np.int64)
    for i in range(n):
        for j in range(i, n):
            A[i, j] = A[j, i] = int(rng.integers(lo, hi + 1))
    return A


def _det_nonzero(L):
    """Exact det(L) != 0 via fraction-free elimination over Python ints."""
    M = [[int(v) for v in row] for row in L]
    n = len(M)
    prev = 1
    for k in range(n):
        piv = next((r for r in range(k, n) if M[r][k] != 0), None)
        if piv is None:
            return False
        M[k], M[piv] = M[piv], M[k]
        for r in range(k + 1, n):
            for c in range(k + 1, n):
                M[r][c] = (M[r][c] * M[k][k] - M[r][k] * M[k][c]) // prev
        prev = M[k][k]
    return M[n - 1][n - 1] != 0


def lattice(m=24, d=24, lo=-9, hi=9, seed=1):
    rng = np.random.default_rng(seed)
    while True:
        L = rng.integers(lo, hi + 1, (m, d)).astype(np.int64)
        if m > d or _det_nonzero(L if m == d else L @ L.T):
            return L, (L @ L.T).astype(np.int64)


def planted(n, dmax=100, shears=None, seed=1):
    """Returns (A, D, U) with A = U^T D U exactly."""
    shears = 3 * n if shears is None else shears
    rng = np.random.default_rng(seed)
    while True:
        d = rng.integers(1, dmax + 1, n) * np.where(rng.integers(0, 2, n) == 1, 1, -1)
        U = np.eye(n, dtype=np.int64)
        for _ in range(shears):
            i = int(rng.integers(0, n))
            j = int(rng.integers(0, n - 1))
            j += j >= i
            s = 1 if rng.integers(0, 2) else -1
            U[:, j] += s * U[:, i]
        mu = int(np.abs(U).max())
        if n * mu * mu * dmax < (1 << 62):
            A = (U.T * d) @ U
        else:  # exact fallback
            Uo = U.astype(object)
            A = (Uo.T * d.astype(object)).dot(Uo)
        if max(abs(int(v)) for v in np.asarray(A).ravel()) <= LIMIT:
            return np.asarray(A, dtype=np.int64), np.diag(d).astype(np.int64), U


def batch_config(count=1024, nmin=8, nmax=48, dmax=100, seed=7):
    """The batching workload: `count` planted instances with n ~ U{nmin..nmax}."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, count)
    return [planted(int(n), dmax, 3 * int(n), seed * 1_000_003 + q)[0] for q, n in enumerate(ns)]


def serve_config(count=30000, dmax=200, seed=11, rate=None):
    """The serving workload: (arrival_time_s, A) with n log-uniform in [6, 64]
    and Poisson arrivals at `rate` per second (None: all at t = 0)."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(6), math.log(64), count))).astype(int)
    t = np.cumsum(rng.exponential(1.0 / rate, count)) if rate else np.zeros(count)
    return [(float(t[q]), planted(int(ns[q]), dmax, 3 * int(ns[q]), seed * 1_000_003 + q)[0])
            for q in range(count)]
