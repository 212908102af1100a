"""ctypes wrapper around oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [
        ("t_init", C.c_double), ("t_min", C.c_double), ("cooling", C.c_double),
        ("reheat", C.c_double), ("reseed_factor", C.c_double),
        ("greedy_fraction", C.c_double), ("target_fraction", C.c_double),
        ("target_accept", C.c_double), ("worker_diversity", C.c_double),
        ("sweep_fraction", C.c_double), ("moves_per_cool", C.c_int32),
        ("stuck_threshold", C.c_int32), ("target_samples", C.c_int32),
        ("tempering", C.c_int32), ("adaptive_cooling", C.c_int32),
        ("exchange_interval", C.c_int32), ("replicas", C.c_int32),
        ("reserved", C.c_int32), ("seed", C.c_uint64), ("max_moves", C.c_uint64),
    ]


class VerifyReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("primes_used", C.c_int32),
                ("det_sign", C.c_int32), ("reserved", C.c_int32)]


class ModelResult(C.Structure):
    _fields_ = [
        ("solved", C.c_int32), ("solve_replica", C.c_int32), ("solve_move", C.c_int64),
        ("best_score", C.c_int64), ("best_replica", C.c_int32), ("epochs", C.c_int32),
        ("total_moves", C.c_int64), ("accepts", C.c_int64), ("exchanges", C.c_int64),
    ]


class DualParams(C.Structure):
    _fields_ = [("c", C.c_double), ("dual_tol", C.c_double), ("dual_refresh", C.c_int32),
                ("dual_check", C.c_int32)]


class DualRefResult(C.Structure):
    _fields_ = [
        ("solved", C.c_int32), ("reserved", C.c_int32), ("moves", C.c_int64),
        ("accepts", C.c_int64), ("refreshes", C.c_int64), ("checks", C.c_int64),
        ("recomputes", C.c_int64), ("score", C.c_double), ("gb_score", C.c_double),
        ("c", C.c_double),
    ]


class DualResult(C.Structure):
    _fields_ = [
        ("solved", C.c_int32), ("solve_replica", C.c_int32), ("solve_move", C.c_int64),
        ("best_score", C.c_double), ("best_replica", C.c_int32), ("epochs", C.c_int32),
        ("total_moves", C.c_int64), ("accepts", C.c_int64), ("exchanges", C.c_int64),
        ("refreshes", C.c_int64), ("checks", C.c_int64), ("recomputes", C.c_int64),
    ]


class RefParams(C.Structure):
    _fields_ = [("base", Params), ("worker_seed", C.c_uint32), ("thread_id", C.c_int32),
                ("rung_T", C.c_double)]


class RefResult(C.Structure):
    _fields_ = [("solved", C.c_int32), ("reserved", C.c_int32), ("moves", C.c_int64),
                ("accepts", C.c_int64), ("score", C.c_int64), ("gb_score", C.c_int64)]


LOG_DTYPE = np.dtype([("move", "<i8"), ("i", "<i2"), ("j", "<i2"), ("s", "<i2"),
                      ("flags", "<u2")])

# Reference defaults (congruence_anneal.hpp:122-153).
DEFAULTS = dict(t_init=0.0, t_min=1e-3, cooling=0.999, reheat=1.0, reseed_factor=1.25,
                greedy_fraction=0.5, target_fraction=0.5, target_accept=0.44,
                worker_diversity=0.4, sweep_fraction=0.0, moves_per_cool=200,
                stuck_threshold=20000, target_samples=8, tempering=1, adaptive_cooling=0,
                exchange_interval=2000, replicas=8, seed=1, max_moves=0)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_rounded_div.restype = C.c_int64
        L.orc_rounded_div.argtypes = [C.c_int64, C.c_int64]
        L.orc_score.restype = C.c_int64
        L.orc_score.argtypes = [I64P, C.c_int32]
        L.orc_offdiag_nonzero.restype = C.c_int64
        L.orc_offdiag_nonzero.argtypes = [I64P, C.c_int32]
        L.orc_try_add.restype = C.c_int32
        L.orc_try_add.argtypes = [I64P, I64P, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                                  I64P, I64P, I64P, I64P]
        L.orc_add_transform_row.restype = C.c_int32
        L.orc_add_transform_row.argtypes = [I64P, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_int64, I64P]
        L.orc_commit_add.restype = None
        L.orc_commit_add.argtypes = [I64P, I64P, C.c_int32, C.c_int32, I64P, I64P, I64P]
        L.orc_compute_row_mass.restype = None
        L.orc_compute_row_mass.argtypes = [I64P, C.c_int32, I64P]
        L.orc_score_delta_l1.restype = C.c_int64
        L.orc_score_delta_l1.argtypes = [I64P, C.c_int32, C.c_int32, C.c_int32, C.c_int64]
        L.orc_best_shear_l1.restype = C.c_int64
        L.orc_best_shear_l1.argtypes = [I64P, C.c_int32, C.c_int32, C.c_int32]
        L.orc_congruence_of.restype = None
        L.orc_congruence_of.argtypes = [I64P, I64P, C.c_int32, I64P]
        L.orc_gram_of.restype = None
        L.orc_gram_of.argtypes = [I64P, C.c_int32, C.c_int32, I64P]
        L.orc_xt_times.restype = None
        L.orc_xt_times.argtypes = [I64P, I64P, C.c_int32, C.c_int32, I64P]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_xexp.restype = C.c_double
        L.orc_xexp.argtypes = [C.c_double]
        L.orc_verify.restype = C.c_int32
        L.orc_verify.argtypes = [I64P, I64P, I64P, C.c_int32, C.POINTER(VerifyReport)]
        L.orc_replay.restype = C.c_int64
        L.orc_replay.argtypes = [I64P, I64P, C.c_int32, C.c_void_p, C.c_int64]
        L.orc_model_run.restype = C.c_int32
        L.orc_model_run.argtypes = [I64P, I64P, C.c_int32, I32P, C.c_int32,
                                    C.POINTER(Params), C.c_int32, I64P, I64P, I64P, I64P,
                                    I64P, C.c_void_p, I64P, C.c_int64,
                                    C.POINTER(ModelResult)]
        L.orc_ladder.restype = None
        L.orc_ladder.argtypes = [C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_double)]
        L.orc_fractions.restype = None
        L.orc_fractions.argtypes = [C.POINTER(Params), C.c_int32, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.orc_suggest_t_init.restype = C.c_double
        L.orc_suggest_t_init.argtypes = [I64P, C.c_int32]
        DP = C.POINTER(C.c_double)
        L.orc_pairsum.restype = C.c_double
        L.orc_pairsum.argtypes = [DP, C.c_int32]
        L.orc_dual_eval.restype = C.c_int32
        L.orc_dual_eval.argtypes = [I64P, DP, C.c_int32, C.c_int32, C.c_int32, C.c_int64, I64P,
                                    DP, DP, DP, I64P]
        L.orc_dual_best_shear.restype = C.c_int64
        L.orc_dual_best_shear.argtypes = [I64P, DP, C.c_int32, C.c_double, C.c_int32, C.c_int32]
        L.orc_invert.restype = C.c_int32
        L.orc_invert.argtypes = [I64P, C.c_int32, DP]
        L.orc_frob2_off_p.restype = C.c_double
        L.orc_frob2_off_p.argtypes = [I64P, C.c_int32]
        L.orc_frob2_off_q.restype = C.c_double
        L.orc_frob2_off_q.argtypes = [DP, C.c_int32]
        L.orc_dual_score.restype = C.c_double
        L.orc_dual_score.argtypes = [I64P, DP, C.c_int32, C.c_double]
        VP = C.c_void_p
        L.orc_dual_model_run.restype = C.c_int32
        L.orc_dual_model_run.argtypes = [I64P, DP, I64P, C.c_int32, VP, C.c_int32, VP, VP,
                                         C.c_int32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                         C.c_int64, VP]
        L.orc_exchange_sweep.restype = C.c_int64
        L.orc_exchange_sweep.argtypes = [DP, DP, I32P, C.c_int32, C.c_uint64, C.c_uint64]
        L.orc_set_mutation.restype = None
        L.orc_set_mutation.argtypes = [C.c_int32]
        L.orc_get_mutation.restype = C.c_int32
        L.orc_get_mutation.argtypes = []
        L.orc_metropolis.restype = C.c_int32
        L.orc_metropolis.argtypes = [C.c_int64, C.c_double, C.c_uint32, C.c_uint32]
        L.orc_ref_worker_seed.restype = C.c_uint32
        L.orc_ref_worker_seed.argtypes = [C.c_uint64, C.c_int32]
        L.orc_ref_worker_run.restype = C.c_int32
        L.orc_ref_worker_run.argtypes = [I64P, I64P, C.c_int32, VP, C.c_int32,
                                         C.POINTER(RefParams), VP, VP, VP, VP, VP, VP,
                                         C.c_int64, C.POINTER(RefResult)]
        L.orc_dual_ref_run.restype = C.c_int32
        L.orc_dual_ref_run.argtypes = [I64P, I64P, C.c_int32, VP, C.c_int32, VP, VP, C.c_double,
                                       VP, VP, VP, VP, VP, VP, VP, VP, C.c_int64, VP]
        L.orc_dual_residual.restype = C.c_double
        L.orc_dual_residual.argtypes = [I64P, DP, C.c_int32, C.c_int32]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(I64P)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def rounded_div(a, b):
    return lib().orc_rounded_div(int(a), int(b))


def score(A):
    A = _i64(A)
    return lib().orc_score(_p(A), A.shape[0])


def offdiag_nonzero(A):
    A = _i64(A)
    return lib().orc_offdiag_nonzero(_p(A), A.shape[0])


def try_add(A, Xt, i, j, s, with_xt=False):
    A = _i64(A)
    Xt = _i64(Xt)
    n = A.shape[0]
    row = np.zeros(n, np.int64)
    xrow = np.zeros(n, np.int64)
    d = C.c_int64(0)
    dnz = C.c_int64(0)
    ok = lib().orc_try_add(_p(A), _p(Xt), n, i, j, s, _p(row),
                           _p(xrow) if with_xt else None, C.byref(d), C.byref(dnz))
    return bool(ok), d.value, dnz.value, row, xrow


def add_transform_row(Xt, i, j, s):
    Xt = _i64(Xt)
    out = np.zeros(Xt.shape[0], np.int64)
    ok = lib().orc_add_transform_row(_p(Xt), Xt.shape[0], i, j, s, _p(out))
    return bool(ok), out


def commit_add(A, Xt, j, row, xrow, mass):
    A = _i64(A).copy()
    Xt = _i64(Xt).copy()
    mass = _i64(mass).copy()
    lib().orc_commit_add(_p(A), _p(Xt), A.shape[0], j, _p(_i64(row)), _p(_i64(xrow)),
                         _p(mass))
    return A, Xt, mass


def compute_row_mass(A):
    A = _i64(A)
    out = np.zeros(A.shape[0], np.int64)
    lib().orc_compute_row_mass(_p(A), A.shape[0], _p(out))
    return out


def score_delta_l1(A, i, j, s):
    A = _i64(A)
    return lib().orc_score_delta_l1(_p(A), A.shape[0], i, j, s)


def best_shear_l1(A, i, j):
    A = _i64(A)
    return lib().orc_best_shear_l1(_p(A), A.shape[0], i, j)


def congruence_of(A, X):
    A = _i64(A)
    X = _i64(X)
    out = np.zeros_like(A)
    lib().orc_congruence_of(_p(A), _p(X), A.shape[0], _p(out))
    return out


def gram_of(L):
    L = _i64(L)
    m, d = L.shape
    out = np.zeros((m, m), np.int64)
    lib().orc_gram_of(_p(L), m, d, _p(out))
    return out


def xt_times(X, L):
    X = _i64(X)
    L = _i64(L)
    m, d = L.shape
    out = np.zeros((m, d), np.int64)
    lib().orc_xt_times(_p(X), _p(L), m, d, _p(out))
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[t] for t in range(4)]


def xexp(x):
    return lib().orc_xexp(float(x))


def verify(A0, X, A):
    A0 = _i64(A0)
    X = _i64(X)
    A = _i64(A)
    rep = VerifyReport()
    lib().orc_verify(_p(A0), _p(X), _p(A), A0.shape[0], C.byref(rep))
    return rep.status, rep.det_sign, rep.primes_used


def replay(A0, Xt0, log):
    A = _i64(A0).copy()
    Xt = _i64(Xt0).copy()
    log = np.ascontiguousarray(log, dtype=LOG_DTYPE)
    k = lib().orc_replay(_p(A), _p(Xt), A.shape[0], log.ctypes.data, len(log))
    return k, A, Xt


def make_params(**kw):
    p = Params()
    vals = dict(DEFAULTS)
    vals.update(kw)
    for k, v in vals.items():
        setattr(p, k, v)
    return p


def ladder(top, t_min, R):
    out = (C.c_double * R)()
    lib().orc_ladder(top, t_min, R, out)
    return [out[k] for k in range(R)]


def fractions(params, r):
    g = C.c_double()
    t = C.c_double()
    lib().orc_fractions(C.byref(params), r, C.byref(g), C.byref(t))
    return g.value, t.value


def suggest_t_init(A):
    A = _i64(A)
    return lib().orc_suggest_t_init(_p(A), A.shape[0])


def model_run(A0, X0, params, max_epochs, active=None, log_cap=0):
    """Runs the trajectory model; returns a dict of per-replica arrays."""
    A0 = _i64(A0)
    n = A0.shape[0]
    Xt0 = _i64(np.asarray(X0).T)
    R = params.replicas
    A_out = np.zeros((R, n, n), np.int64)
    Xt_out = np.zeros((R, n, n), np.int64)
    bA = np.zeros((R, n, n), np.int64)
    bXt = np.zeros((R, n, n), np.int64)
    scal = np.zeros((R, 8), np.int64)
    logs = np.zeros(max(1, R * log_cap), LOG_DTYPE)
    counts = np.zeros(R, np.int64)
    res = ModelResult()
    act = None
    na = 0
    if active is not None:
        act = np.ascontiguousarray(active, dtype=np.int32)
        na = len(act)
    lib().orc_model_run(_p(A0), _p(Xt0), n,
                        act.ctypes.data_as(I32P) if act is not None else None, na,
                        C.byref(params), max_epochs, _p(A_out), _p(Xt_out), _p(bA), _p(bXt),
                        _p(scal), logs.ctypes.data if log_cap else None,
                        _p(counts) if log_cap else None, log_cap, C.byref(res))
    out = dict(A=A_out, Xt=Xt_out, best_A=bA, best_Xt=bXt, scal=scal,
               solved=res.solved, solve_replica=res.solve_replica,
               solve_move=res.solve_move, best_score=res.best_score,
               best_replica=res.best_replica, epochs=res.epochs,
               total_moves=res.total_moves, accepts=res.accepts, exchanges=res.exchanges)
    if log_cap:
        out["logs"] = [logs[r * log_cap: r * log_cap + counts[r]].copy() for r in range(R)]
    return out


def exchange_sweep(temps, scores, rungs, seed, epoch):
    """The model's exchange sweep; returns (new rungs, swaps)."""
    temps = np.ascontiguousarray(temps, dtype=np.float64)
    scores = np.ascontiguousarray(scores, dtype=np.float64)
    r = np.ascontiguousarray(rungs, dtype=np.int32).copy()
    DP = C.POINTER(C.c_double)
    acc = lib().orc_exchange_sweep(temps.ctypes.data_as(DP), scores.ctypes.data_as(DP),
                                   r.ctypes.data_as(I32P), len(r), int(seed), int(epoch))
    return r, acc


def set_mutation(which):
    """Plants a mistake in the model (0 = none); see oracle.h."""
    lib().orc_set_mutation(int(which))


def metropolis(d, T, u_hi, u_lo):
    return lib().orc_metropolis(int(d), float(T), int(u_hi), int(u_lo))


def ref_worker_seed(seed, t):
    """Worker t's seed in a reference run with --seed `seed`
    (congruence_anneal.hpp:484-488)."""
    return lib().orc_ref_worker_seed(int(seed), int(t))


def ref_worker_run(A0, X0, worker_seed, threads=1, thread_id=0, rung_T=0.0, active=None,
                   log_cap=0, **kw):
    """The model's loop body with the reference's draws and publish rule for
    one worker (see oracle.h). kw: engine parameters (tempering, t_init, ...,
    max_moves). Returns a dict: A, Xt (final), gbA, gbXt (global best), log,
    solved, moves, accepts, score, gb_score."""
    A0 = _i64(A0)
    n = A0.shape[0]
    Xt0 = _i64(np.asarray(X0).T)
    rp = RefParams()
    rp.base = make_params(**{**kw, "replicas": threads})
    rp.worker_seed = int(worker_seed) & 0xFFFFFFFF
    rp.thread_id = thread_id
    rp.rung_T = rung_T
    A = np.zeros((n, n), np.int64)
    Xt = np.zeros((n, n), np.int64)
    gA = np.zeros((n, n), np.int64)
    gXt = np.zeros((n, n), np.int64)
    log = np.zeros(max(1, log_cap), LOG_DTYPE)
    cnt = np.zeros(1, np.int64)
    act = np.ascontiguousarray(active, dtype=np.int32) if active is not None else None
    res = RefResult()
    lib().orc_ref_worker_run(_p(A0), _p(Xt0), n, act.ctypes.data if act is not None else None,
                             len(act) if act is not None else 0, C.byref(rp), A.ctypes.data,
                             Xt.ctypes.data, gA.ctypes.data, gXt.ctypes.data,
                             log.ctypes.data if log_cap else None,
                             cnt.ctypes.data if log_cap else None, log_cap, C.byref(res))
    return dict(A=A, Xt=Xt, gbA=gA, gbXt=gXt, log=log[:cnt[0]].copy() if log_cap else None,
                solved=res.solved, moves=res.moves, accepts=res.accepts, score=res.score,
                gb_score=res.gb_score)


# ---- xdual primal/dual objective (oracle_dual.cpp) ----
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def pairsum(x):
    x = _f64(x)
    return lib().orc_pairsum(_pd(x), len(x))


def dual_eval(P, Q, i, j, s):
    """(feasible, d_off_p, d_off_q, d_nonzero, new P row j, new Q row i)."""
    P, Q = _i64(P), _f64(Q)
    n = P.shape[0]
    prow, qrow = np.zeros(n, np.int64), np.zeros(n)
    dp, dq, dnz = C.c_double(), C.c_double(), C.c_int64()
    ok = lib().orc_dual_eval(_p(P), _pd(Q), n, i, j, s, _p(prow), _pd(qrow), C.byref(dp),
                             C.byref(dq), C.byref(dnz))
    return bool(ok), dp.value, dq.value, dnz.value, prow, qrow


def dual_best_shear(P, Q, c, i, j):
    P, Q = _i64(P), _f64(Q)
    return lib().orc_dual_best_shear(_p(P), _pd(Q), P.shape[0], float(c), i, j)


def invert(P):
    P = _i64(P)
    Q = np.zeros(P.shape)
    ok = lib().orc_invert(_p(P), P.shape[0], _pd(Q))
    return bool(ok), Q


def frob2_off_p(P):
    P = _i64(P)
    return lib().orc_frob2_off_p(_p(P), P.shape[0])


def frob2_off_q(Q):
    Q = _f64(Q)
    return lib().orc_frob2_off_q(_pd(Q), Q.shape[0])


def dual_residual(P, Q, k):
    P, Q = _i64(P), _f64(Q)
    return lib().orc_dual_residual(_p(P), _pd(Q), P.shape[0], k)


def dual_score(P, Q, c):
    P, Q = _i64(P), _f64(Q)
    return lib().orc_dual_score(_p(P), _pd(Q), P.shape[0], float(c))


def dual_model_run(P0, Q0, X0, params, dparams, max_epochs, active=None, log_cap=0):
    """Runs the engine model on the primal/dual objective from (P0, Q0, X0)."""
    P0, Q0 = _i64(P0), _f64(Q0)
    n = P0.shape[0]
    Xt0 = _i64(np.asarray(X0).T)
    R = params.replicas
    P = np.zeros((R, n, n), np.int64)
    Q = np.zeros((R, n, n))
    Xt = np.zeros((R, n, n), np.int64)
    dscal = np.zeros((R, 4))
    scal = np.zeros((R, 8), np.int64)
    bP = np.zeros((n, n), np.int64)
    bQ = np.zeros((n, n))
    bXt = np.zeros((n, n), np.int64)
    logs = np.zeros(max(1, R * log_cap), LOG_DTYPE)
    counts = np.zeros(R, np.int64)
    res = DualResult()
    act = None
    na = 0
    if active is not None:
        act = np.ascontiguousarray(active, dtype=np.int32)
        na = len(act)
    d = lambda a: a.ctypes.data  # noqa: E731
    lib().orc_dual_model_run(_p(P0), _pd(Q0), _p(Xt0), n, d(act) if act is not None else None,
                             na, C.addressof(params), C.addressof(dparams), max_epochs, d(P),
                             d(Q), d(Xt), d(dscal), d(scal), d(bP), d(bQ), d(bXt),
                             d(logs) if log_cap else None, d(counts) if log_cap else None,
                             log_cap, C.addressof(res))
    out = dict(P=P, Q=Q, Xt=Xt, dscal=dscal, scal=scal, best_P=bP, best_Q=bQ, best_Xt=bXt,
               **{f: getattr(res, f) for f, _ in DualResult._fields_})
    if log_cap:
        out["logs"] = [logs[r * log_cap: r * log_cap + counts[r]].copy() for r in range(R)]
    return out


def dual_ref_run(P0, X0, worker_seed, threads=1, thread_id=0, rung_T=0.0, lam=1.0, active=None,
                 dual_refresh=200000, dual_check=25000, dual_tol=1e-6, log_cap=0, **kw):
    """The dual model's loop body in the reference's mode: one worker of
    xdual's annealer from (P0, X0) (see oracle.h). kw: engine parameters
    (tempering, t_init, ..., max_moves). Returns a dict: P, Q, Xt (the
    worker's final state), gbP, gbQ, gbXt (the global best), log, and the
    DualRefResult fields."""
    P0 = _i64(P0)
    n = P0.shape[0]
    Xt0 = _i64(np.asarray(X0).T)
    rp = RefParams()
    rp.base = make_params(**{**kw, "replicas": threads})
    rp.worker_seed = int(worker_seed) & 0xFFFFFFFF
    rp.thread_id = thread_id
    rp.rung_T = rung_T
    dp = DualParams()
    dp.c, dp.dual_tol, dp.dual_refresh, dp.dual_check = 0.0, dual_tol, dual_refresh, dual_check
    P, gP = np.zeros((n, n), np.int64), np.zeros((n, n), np.int64)
    Q, gQ = np.zeros((n, n)), np.zeros((n, n))
    Xt, gXt = np.zeros((n, n), np.int64), np.zeros((n, n), np.int64)
    log = np.zeros(max(1, log_cap), LOG_DTYPE)
    cnt = np.zeros(1, np.int64)
    act = np.ascontiguousarray(active, dtype=np.int32) if active is not None else None
    res = DualRefResult()
    d = lambda a: a.ctypes.data  # noqa: E731
    lib().orc_dual_ref_run(_p(P0), _p(Xt0), n, d(act) if act is not None else None,
                           len(act) if act is not None else 0, C.addressof(rp), C.addressof(dp),
                           float(lam), d(P), d(Q), d(Xt), d(gP), d(gQ), d(gXt),
                           d(log) if log_cap else None, d(cnt) if log_cap else None, log_cap,
                           C.addressof(res))
    out = dict(P=gP, Q=gQ, Xt=gXt, Pf=P, Qf=Q, Xtf=Xt, log=log[:cnt[0]].copy() if log_cap else None)
    out.update({f: getattr(res, f) for f, _ in DualRefResult._fields_})
    return out
