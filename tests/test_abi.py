"""CPU checks of the product boundary: libxtx.so loads, exports exactly what
include/xtx.h declares, and its host-only parts (defaults, verifier) behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xtx.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(xtx_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ["xtx_open", "xtx_close", "xtx_last_error", "xtx_query_limits",
                 "xtx_solve_plain", "xtx_batch_solve", "xtx_serve_submit", "xtx_serve_poll",
                 "xtx_serve_retire", "xtx_movelog_get", "xtx_movelog_free", "xtx_verify"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from xtax_amd import _native
    lib = _native.lib()
    for name in declared():
        assert hasattr(lib, name), name
    assert set(declared()) == set(_native.SIGNATURES)


def test_params_defaults_mirror_engine_params():
    # canneal::EngineParams defaults (congruence_anneal.hpp:122-153)
    from xtax_amd import EngineParams, _native
    p = _native.Params()
    _native.lib().xtx_params_default(C.byref(p))
    want = dict(t_init=0.0, t_min=1e-3, cooling=0.999, moves_per_cool=200, reheat=1.0,
                stuck_threshold=20000, reseed_factor=1.25, greedy_fraction=0.5,
                target_fraction=0.5, target_samples=8, max_seconds=0.0, save_interval=2.0,
                tempering=1, exchange_interval=2000, adaptive_cooling=0, target_accept=0.44,
                worker_diversity=0.4, sweep_fraction=0.0)
    for k, v in want.items():
        assert getattr(p, k) == v, k
        assert getattr(EngineParams(), k) == v, k


def test_open_without_gpu_fails_cleanly():
    from xtax_amd import _native
    lib = _native.lib()
    h = C.c_void_p()
    rc = lib.xtx_open(0, C.byref(h))
    if rc == 0:
        lib.xtx_close(h)
        pytest.skip("a GPU is present")
    assert rc < 0 and lib.xtx_last_error(None)


def test_verifier_matches_oracle(orc):
    from xtax_amd import verify
    g = np.load(os.path.join(ROOT, "tests", "golden", "g2_end_to_end.npz"))
    for t, solved, moves, det in g["meta"].tolist():
        A0, A, X = g[f"r{t}_A0"], g[f"r{t}_A"], g[f"r{t}_X"]
        st, sign = verify(A0, X, A)
        ost, osign, _ = orc.verify(A0, X, A)
        assert st == ost and sign == osign
    A0, A, X = g["r0_A0"], g["r0_A"], g["r0_X"]
    Xb = X.copy()
    Xb[1, 2] += 1
    assert verify(A0, Xb, A)[0] & 1
    n = A0.shape[0]
    X2 = np.eye(n, dtype=np.int64)
    X2[0, 0] = -2
    st, _ = verify(A0, X2, X2.T @ A0 @ X2)
    assert st & 4 and not st & 1


def test_verifier_large_entries_modular_path(orc):
    """Entries near 2^48 force the multi-modular path; planted mistakes in the
    low and high bits are both caught."""
    from xtax_amd import verify
    rng = np.random.default_rng(4)
    n = 5
    X = np.eye(n, dtype=np.int64)
    for _ in range(30):
        i, j = rng.choice(n, 2, replace=False)
        X[:, j] += int(rng.choice([-3, 3])) * X[:, i]
        if np.abs(X).max() > (1 << 30):
            break
    d = [1, -1, 2, 3, -5]
    Xi = [[int(v) for v in r] for r in X.tolist()]
    A = [[sum(Xi[k][r] * d[k] * Xi[k][c] for k in range(n)) for c in range(n)] for r in range(n)]
    if max(abs(v) for r in A for v in r) <= (1 << 62):
        A = np.array(A, np.int64)
        A0 = np.diag(d).astype(np.int64)
        st, sign = verify(A0, X, A)
        assert st & 1 == 0 and st & 4 == 0
        assert (st, sign) == orc.verify(A0, X, A)[:2]
        for bit in (0, 40):
            Ab = A.copy()
            Ab[0, 0] += 1 << bit
            assert verify(A0, X, Ab)[0] & 1


def test_workload_generators_are_exact():
    from xtax_amd import workloads
    A, D, U = workloads.planted(12, 100, 36, seed=3)
    Uo = U.astype(object)
    assert (Uo.T.dot(D.astype(object)).dot(Uo) == A.astype(object)).all()
    assert (A == A.T).all()
    L, G = workloads.lattice(6, 6, seed=2)
    assert (G == L @ L.T).all()
    R = workloads.random_symmetric(10, seed=5)
    assert (R == R.T).all() and R.min() >= -20 and R.max() <= 20
    batch = workloads.batch_config(count=5)
    assert all(8 <= a.shape[0] <= 48 for a in batch)
