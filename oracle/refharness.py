"""ctypes wrapper of the reference's own worker loop, recorded
(oracle/_ref/libref_harness.so: ref_worker_trace on the L1 objective;
oracle/_ref/libref_dual.so: refd_worker_trace on xdual's objective; built from
the reference sources by `make -C oracle ref`).

TEST INFRASTRUCTURE ONLY: used by tests/golden/gen_golden.py to record worker
traces and by the tests, where the library exists, to compare the oracle
model's reference-draw mode with canneal::anneal_worker
(congruence_anneal.hpp:177-416) move for move.
"""
import ctypes as C
import os

import numpy as np

from . import pyoracle as orc

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libref_harness.so")
DUAL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libref_dual.so")
_LIB = None
_DLIB = None

# the order of the harness's parameter arrays (oracle/ref_harness.cpp)
DKEYS = ("t_init", "t_min", "cooling", "reheat", "reseed_factor", "greedy_fraction",
         "target_fraction", "target_accept", "worker_diversity", "sweep_fraction")
IKEYS = ("moves_per_cool", "stuck_threshold", "target_samples", "tempering", "adaptive_cooling")


def available():
    return os.path.exists(PATH)


def dual_available():
    return os.path.exists(DUAL_PATH)


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(PATH)
        P = C.c_void_p
        L.ref_worker_trace.restype = C.c_int
        L.ref_worker_trace.argtypes = [P, P, C.c_int, P, C.c_int, P, P, P, P, P, P, P, C.c_int64,
                                       P]
        _LIB = L
    return _LIB


def pack(kw, threads, thread_id, seed, cap, rung_T):
    """(dp, ip) parameter arrays of one trace."""
    d = {**orc.DEFAULTS, **kw}
    dp = np.array([d[k] for k in DKEYS] + [rung_T], np.float64)
    ip = np.array([d[k] for k in IKEYS] + [threads, thread_id, seed, cap], np.int64)
    return dp, ip


def unpack(dp, ip):
    """Inverse of pack: (kw, threads, thread_id, seed, cap, rung_T)."""
    kw = {k: float(dp[q]) for q, k in enumerate(DKEYS)}
    kw.update({k: int(ip[q]) for q, k in enumerate(IKEYS)})
    return kw, int(ip[5]), int(ip[6]), int(ip[7]), int(ip[8]), float(dp[10])


def trace(A_start, X0, dp, ip, active=None, trace_cap=400000):
    """One worker of the reference, capped at ip[8] moves. Returns a dict:
    A, Xt (final), gbA, gbXt (global best), log (engine log format), moves,
    score, gb_score, solved."""
    A_start = np.ascontiguousarray(A_start, np.int64)
    X0 = np.ascontiguousarray(X0, np.int64)
    n = A_start.shape[0]
    dp = np.ascontiguousarray(dp, np.float64)
    ip = np.ascontiguousarray(ip, np.int64)
    A = np.zeros((n, n), np.int64)
    Xt, gA, gXt = A.copy(), A.copy(), A.copy()
    tr = np.zeros((trace_cap, 4), np.int64)
    sc = np.zeros(5, np.int64)
    act = np.ascontiguousarray(active, np.int32) if active is not None and len(active) else None
    rc = lib().ref_worker_trace(A_start.ctypes.data, X0.ctypes.data, n,
                                act.ctypes.data if act is not None else None,
                                len(act) if act is not None else 0, dp.ctypes.data,
                                ip.ctypes.data, A.ctypes.data, Xt.ctypes.data, gA.ctypes.data,
                                gXt.ctypes.data, tr.ctypes.data, trace_cap, sc.ctypes.data)
    if rc != 0:
        raise ValueError("ref_worker_trace needs a positive move cap")
    tr = tr[:min(int(sc[4]), trace_cap)]
    log = np.zeros(len(tr), orc.LOG_DTYPE)
    log["move"] = tr[:, 0]
    log["i"] = tr[:, 1]
    log["j"] = tr[:, 2]
    log["s"] = (tr[:, 3] & 0xFFFF).astype(np.uint16).view(np.int16)
    log["flags"] = tr[:, 3] >> 32
    return dict(A=A, Xt=Xt, gbA=gA, gbXt=gXt, log=log, moves=int(sc[0]), score=int(sc[1]),
                gb_score=int(sc[2]), solved=int(sc[3]))


def model(A_start, X0, dp, ip, active=None, log_cap=400000):
    """The oracle model's reference-draw mode on the same inputs."""
    kw, threads, tid, seed, cap, rung_T = unpack(dp, ip)
    return orc.ref_worker_run(A_start, X0, seed, threads=threads, thread_id=tid, rung_T=rung_T,
                              active=None if active is None or not len(active) else active,
                              log_cap=log_cap, **{**kw, "max_moves": cap})


def dual_lib():
    global _DLIB
    if _DLIB is None:
        L = C.CDLL(DUAL_PATH)
        P = C.c_void_p
        L.refd_worker_trace.restype = C.c_int
        L.refd_worker_trace.argtypes = [P, P, C.c_int, P, C.c_int, P, P, P, P, P, P, P, P, P,
                                        C.c_int64, P, P]
        _DLIB = L
    return _DLIB


def dual_pack(kw, threads, thread_id, seed, cap, rung_T, lam, dual_refresh, dual_check, dual_tol):
    """(dp, ip) parameter arrays of one dual trace."""
    dp, ip = pack(kw, threads, thread_id, seed, cap, rung_T)
    return (np.concatenate([dp, [lam, dual_tol]]),
            np.concatenate([ip, [dual_refresh, dual_check]]).astype(np.int64))


def dual_unpack(dp, ip):
    """Inverse of dual_pack: (kw, threads, thread_id, seed, cap, rung_T, dual
    keywords)."""
    kw, threads, tid, seed, cap, rung_T = unpack(dp[:11], ip[:9])
    return kw, threads, tid, seed, cap, rung_T, dict(
        lam=float(dp[11]), dual_tol=float(dp[12]), dual_refresh=int(ip[9]),
        dual_check=int(ip[10]))


def _log(tr):
    log = np.zeros(len(tr), orc.LOG_DTYPE)
    log["move"] = tr[:, 0]
    log["i"] = tr[:, 1]
    log["j"] = tr[:, 2]
    log["s"] = (tr[:, 3] & 0xFFFF).astype(np.uint16).view(np.int16)
    log["flags"] = tr[:, 3] >> 32
    return log


def dual_trace(P_start, X0, dp, ip, active=None, trace_cap=400000):
    """One worker of the reference on xdual's objective, capped at ip[8]
    moves. Returns a dict: Pf, Qf, Xtf (final), P, Q, Xt (global best), log
    (engine log format), moves, solved, score, gb_score, c."""
    P_start = np.ascontiguousarray(P_start, np.int64)
    X0 = np.ascontiguousarray(X0, np.int64)
    n = P_start.shape[0]
    dp = np.ascontiguousarray(dp, np.float64)
    ip = np.ascontiguousarray(ip, np.int64)
    Pf, gP = np.zeros((n, n), np.int64), np.zeros((n, n), np.int64)
    Qf, gQ = np.zeros((n, n)), np.zeros((n, n))
    Xf, gX = np.zeros((n, n), np.int64), np.zeros((n, n), np.int64)
    tr = np.zeros((trace_cap, 4), np.int64)
    sc = np.zeros(3, np.int64)
    ds = np.zeros(3)
    act = np.ascontiguousarray(active, np.int32) if active is not None and len(active) else None
    d = lambda a: a.ctypes.data  # noqa: E731
    rc = dual_lib().refd_worker_trace(d(P_start), d(X0), n, d(act) if act is not None else None,
                                      len(act) if act is not None else 0, d(dp), d(ip), d(Pf),
                                      d(Qf), d(Xf), d(gP), d(gQ), d(gX), d(tr), trace_cap, d(sc),
                                      d(ds))
    if rc != 0:
        raise ValueError("refd_worker_trace needs a positive move cap and a full-rank start")
    tr = tr[:min(int(sc[2]), trace_cap)]
    return dict(Pf=Pf, Qf=Qf, Xtf=Xf, P=gP, Q=gQ, Xt=gX, log=_log(tr), moves=int(sc[0]),
                solved=int(sc[1]), score=float(ds[0]), gb_score=float(ds[1]), c=float(ds[2]))


def dual_model(P_start, X0, dp, ip, active=None, log_cap=400000):
    """The oracle dual model's reference mode on the same inputs."""
    kw, threads, tid, seed, cap, rung_T, dk = dual_unpack(dp, ip)
    return orc.dual_ref_run(P_start, X0, seed, threads=threads, thread_id=tid, rung_T=rung_T,
                            active=None if active is None or not len(active) else active,
                            log_cap=log_cap, **dk, **{**kw, "max_moves": cap})
