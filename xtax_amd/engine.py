"""Python host API over the gfx950 engine (libxtx.so).

Mirrors the reference's operator interface for this path:
``EngineParams`` has the fields and defaults of ``canneal::EngineParams``
(src/congruence_anneal.hpp:122-153), and ``Engine.run_annealer`` stands where
``canneal::run_annealer<L1Objective>`` does (congruence_anneal.hpp:421-569):
it takes the start matrix and transform and returns the best (or diagonal)
state. Everything runs on the GPU; there is no CPU path.
"""
import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class EngineParams:
    # canneal::EngineParams (congruence_anneal.hpp:122-153)
    t_init: float = 0.0
    t_min: float = 1e-3
    cooling: float = 0.999
    moves_per_cool: int = 200
    reheat: float = 1.0
    stuck_threshold: int = 20000
    reseed_factor: float = 1.25
    greedy_fraction: float = 0.5
    target_fraction: float = 0.5
    target_samples: int = 8
    max_seconds: float = 0.0
    save_interval: float = 2.0
    tempering: bool = True
    exchange_interval: int = 2000
    adaptive_cooling: bool = False
    target_accept: float = 0.44
    worker_diversity: float = 0.4
    sweep_fraction: float = 0.0
    seed: int = 0
    # engine-specific
    replicas: int = 64
    move_log: bool = False
    move_log_cap: int = 0
    max_moves: int = 0
    max_epochs: int = 0
    verify: bool = True
    waves_per_block: int = 0
    autotune: bool = False
    lookahead: int = 0
    persistent: bool = True

    def to_c(self):
        p = N.Params()
        N.lib().xtx_params_default(C.byref(p))
        for f in dataclasses.fields(self):
            setattr(p, f.name, int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                    else getattr(self, f.name))
        return p


@dataclass
class DualParams:
    """xdual's Params beyond the engine's (src/xdual_main.cpp:366-373)."""
    lambda_: float = 1.0
    dual_refresh: int = 200000
    dual_check: int = 25000
    dual_tol: float = 1e-6
    lambda_ramp: float = 0.0

    def to_c(self):
        d = N.DualParams()
        N.lib().xtx_dual_params_default(C.byref(d))
        for f in dataclasses.fields(self):
            setattr(d, f.name, getattr(self, f.name))
        return d


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _p64(a):
    return a.ctypes.data_as(N.I64P) if a is not None else None


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


@dataclass
class SolveResult:
    A: np.ndarray
    X: np.ndarray
    stats: dict = field(default_factory=dict)


@dataclass
class DualResult:
    P: np.ndarray
    Q: np.ndarray
    X: np.ndarray
    stats: dict = field(default_factory=dict)


def _stats(st):
    return {f: getattr(st, f) for f, _ in st._fields_}


class Engine:
    """One device context (xtx_open/xtx_close)."""

    def __init__(self, device=0):
        self._lib = N.lib()
        h = C.c_void_p()
        rc = self._lib.xtx_open(device, C.byref(h))
        if rc != N.XTX_OK:
            raise N.XtxError(rc, self._lib.xtx_last_error(None).decode())
        self._h = h

    def close(self):
        if self._h:
            self._lib.xtx_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != N.XTX_OK:
            raise N.XtxError(rc, self._lib.xtx_last_error(self._h).decode())

    def limits(self):
        L = N.Limits()
        self._check(self._lib.xtx_query_limits(self._h, C.byref(L)))
        d = _stats(L)
        d["arch"] = L.arch.decode()
        return d

    def request_stop(self):
        self._lib.xtx_request_stop(self._h)

    def clear_stop(self):
        self._lib.xtx_clear_stop(self._h)

    # -- plain form -------------------------------------------------------
    def run_annealer(self, A0, X0=None, params=None, active=None, allow_unverified=False,
                     noalloc=False):
        """Anneal X0^T A0 X0 toward a diagonal; returns SolveResult(A, X, stats).

        noalloc: run through xtx_solve_plain_noalloc (the same solve with its
        device state in the library's fixed scratch array)."""
        params = params or EngineParams()
        A0 = _i64(A0)
        n = A0.shape[0]
        X0 = _i64(X0) if X0 is not None else None
        act = np.ascontiguousarray(active, dtype=np.int32) if active is not None else None
        A = np.zeros((n, n), np.int64)
        X = np.zeros((n, n), np.int64)
        st = N.Stats()
        p = params.to_c()
        fn = self._lib.xtx_solve_plain_noalloc if noalloc else self._lib.xtx_solve_plain
        rc = fn(
            self._h, _p64(A0), _p64(X0), n, C.byref(p),
            act.ctypes.data_as(N.I32P) if act is not None else None,
            len(act) if act is not None else 0, _p64(A), _p64(X), C.byref(st))
        if rc == -7 and allow_unverified:
            return SolveResult(A, X, _stats(st))
        self._check(rc)
        return SolveResult(A, X, _stats(st))

    def last_replicas(self, R, n):
        A = np.zeros((R, n, n), np.int64)
        Xt = np.zeros((R, n, n), np.int64)
        sc = np.zeros((R, 8), np.int64)
        self._check(self._lib.xtx_last_replicas(self._h, _p64(A), _p64(Xt), _p64(sc)))
        return A, Xt, sc

    def movelog(self, replica):
        ptr = C.POINTER(N.LogEntry)()
        cnt = C.c_int64()
        tr = C.c_int32()
        self._check(self._lib.xtx_movelog_get(self._h, replica, C.byref(ptr), C.byref(cnt),
                                              C.byref(tr)))
        dt = np.dtype([("move", "<i8"), ("i", "<i2"), ("j", "<i2"), ("s", "<i2"),
                       ("flags", "<u2")])
        if cnt.value == 0:
            return np.zeros(0, dt), bool(tr.value)
        buf = C.string_at(ptr, cnt.value * 16)
        return np.frombuffer(buf, dtype=dt).copy(), bool(tr.value)

    # -- primal/dual objective (xdual) --------------------------------------
    def run_xdual(self, A0, X0=None, params=None, dual=None, active=None,
                  allow_unverified=False):
        """run_xdual (xdual_main.cpp:378-424): anneal P = X0^T A0 X0 and its
        double dual Q = P^-1 under |offdiag P|^2 + c |offdiag Q|^2; returns
        DualResult(P, Q, X, stats)."""
        params = params or EngineParams()
        dual = dual or DualParams()
        A0 = _i64(A0)
        n = A0.shape[0]
        X0 = _i64(X0) if X0 is not None else None
        act = np.ascontiguousarray(active, dtype=np.int32) if active is not None else None
        P = np.zeros((n, n), np.int64)
        Q = np.zeros((n, n), np.float64)
        X = np.zeros((n, n), np.int64)
        st = N.DualStats()
        p, d = params.to_c(), dual.to_c()
        rc = self._lib.xtx_dual_solve(
            self._h, _p64(A0), _p64(X0), n, C.byref(p), C.byref(d),
            act.ctypes.data_as(N.I32P) if act is not None else None,
            len(act) if act is not None else 0, _p64(P), _pd(Q), _p64(X), C.byref(st))
        if rc == -7 and allow_unverified:
            return DualResult(P, Q, X, _stats(st))
        self._check(rc)
        return DualResult(P, Q, X, _stats(st))

    def dual_last_replicas(self, R, n):
        P = np.zeros((R, n, n), np.int64)
        Q = np.zeros((R, n, n), np.float64)
        Xt = np.zeros((R, n, n), np.int64)
        dsc = np.zeros((R, 4), np.float64)
        sc = np.zeros((R, 8), np.int64)
        self._check(self._lib.xtx_dual_last_replicas(self._h, _p64(P), _pd(Q), _p64(Xt),
                                                     _pd(dsc), _p64(sc)))
        return P, Q, Xt, dsc, sc

    def probe_dual(self, P, Q, c, queries):
        """Device primal/dual primitives: (out_i[q, 3] = {feasible, dnz,
        best_shear}, delta[q], P rows[q, n], Q rows[q, n], score[q])."""
        P = _i64(P)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if P.ndim == 2:
            P, Q = P[None], Q[None]
        n = P.shape[1]
        q = np.ascontiguousarray(queries, dtype=np.int32).reshape(-1, 4)
        oi = np.zeros((len(q), 3), np.int64)
        od = np.zeros(len(q))
        rp = np.zeros((len(q), n), np.int64)
        rq = np.zeros((len(q), n))
        sc = np.zeros(len(q))
        self._check(self._lib.xtx_dual_probe(self._h, _p64(P), _pd(Q), n, P.shape[0], float(c),
                                             q.ctypes.data_as(N.I32P), len(q), _p64(oi), _pd(od),
                                             _p64(rp), _pd(rq), _pd(sc)))
        return oi, od, rp, rq, sc

    # -- batched form -----------------------------------------------------
    def batch_solve(self, mats, params=None, seeds=None, replicas=None):
        """Solve many instances of uneven n in one captured launch sequence.

        Returns (list of (A, X, result dict), batch stats dict)."""
        params = params or EngineParams()
        count = len(mats)
        keep = []
        insts = (N.Instance * count)()
        outs = []
        for q, A0 in enumerate(mats):
            A0 = _i64(A0)
            n = A0.shape[0]
            A = np.zeros((n, n), np.int64)
            X = np.zeros((n, n), np.int64)
            keep += [A0, A, X]
            outs.append((A, X))
            insts[q].A0 = _p64(A0)
            insts[q].X0 = None
            insts[q].n = n
            insts[q].replicas = int(replicas[q]) if replicas is not None else 0
            insts[q].seed = int(seeds[q]) if seeds is not None else params.seed + q
            insts[q].active = None
            insts[q].n_active = 0
            insts[q].A_out = _p64(A)
            insts[q].X_out = _p64(X)
        res = (N.Result * count)()
        bs = N.BatchStats()
        p = params.to_c()
        self._check(self._lib.xtx_batch_solve(self._h, insts, count, C.byref(p), res,
                                              C.byref(bs)))
        return [(outs[q][0], outs[q][1], _stats(res[q])) for q in range(count)], _stats(bs)

    # -- launch configuration per order bucket -----------------------------
    def launch_get(self, n):
        out = N.Launch()
        self._check(self._lib.xtx_launch_get(self._h, int(n), C.byref(out)))
        return _stats(out)

    def launch_set(self, order, lanes, cols, triangular, waves_per_block=0, lookahead=1,
                   x_in_lds=0):
        cfg = N.Launch(order, lanes, cols, int(triangular), waves_per_block, 0, 0, lookahead,
                       int(x_in_lds), 0, 0.0)
        self._check(self._lib.xtx_launch_set(self._h, C.byref(cfg)))

    def launch_reset(self, order):
        cfg = N.Launch(order, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0)
        self._check(self._lib.xtx_launch_set(self._h, C.byref(cfg)))

    def autotune(self, mats, params=None, moves=512, seeds=None):
        """Times every launch configuration of each order in `mats` and keeps
        the fastest; returns every measurement (chosen ones with tuned = 1)."""
        params = params or EngineParams()
        count = len(mats)
        insts = (N.Instance * count)()
        keep = []
        for q, A in enumerate(mats):
            A0 = _i64(A)
            keep.append(A0)
            insts[q].A0 = _p64(A0)
            insts[q].X0 = None
            insts[q].n = A0.shape[0]
            insts[q].replicas = 0
            insts[q].seed = int(seeds[q]) if seeds is not None else params.seed + q
            insts[q].active = None
            insts[q].n_active = 0
            insts[q].A_out = None
            insts[q].X_out = None
        cap = 64 * 64
        res = (N.Launch * cap)()
        got = C.c_int32(0)
        p = params.to_c()
        self._check(self._lib.xtx_autotune(self._h, insts, count, C.byref(p), int(moves), res,
                                           cap, C.byref(got)))
        return [_stats(res[q]) for q in range(got.value)]

    # -- served form ------------------------------------------------------
    def serve_open(self, params=None, lds_budget=0, hbm_budget=0, max_resident=0,
                   queue_capacity=0):
        params = params or EngineParams()
        cfg = N.ServeConfig(lds_budget, hbm_budget, max_resident, queue_capacity)
        p = params.to_c()
        self._check(self._lib.xtx_serve_open(self._h, C.byref(p), C.byref(cfg)))
        self._serve_keep = {}

    def serve_submit(self, A0, tag, seed, replicas=0):
        A0 = _i64(A0)
        n = A0.shape[0]
        A = np.zeros((n, n), np.int64)
        X = np.zeros((n, n), np.int64)
        inst = N.Instance()
        inst.A0 = _p64(A0)
        inst.X0 = None
        inst.n = n
        inst.replicas = replicas
        inst.seed = seed
        inst.active = None
        inst.n_active = 0
        inst.A_out = _p64(A)
        inst.X_out = _p64(X)
        self._check(self._lib.xtx_serve_submit(self._h, C.byref(inst), tag))
        self._serve_keep[tag] = (A0, A, X)

    def serve_step(self):
        r = C.c_int32()
        q = C.c_int32()
        self._check(self._lib.xtx_serve_step(self._h, C.byref(r), C.byref(q)))
        return r.value, q.value

    def serve_poll(self, max_results=4096):
        buf = (N.Result * max_results)()
        cnt = C.c_int32()
        self._check(self._lib.xtx_serve_poll(self._h, buf, max_results, C.byref(cnt)))
        out = []
        for t in range(cnt.value):
            d = _stats(buf[t])
            A0, A, X = self._serve_keep.pop(d["tag"])
            out.append((A, X, d))
        return out

    def serve_retire(self, tag):
        self._check(self._lib.xtx_serve_retire(self._h, tag))

    def serve_close(self):
        self._check(self._lib.xtx_serve_close(self._h))

    # -- primitives (parity probes) ---------------------------------------
    def probe_moves(self, A, Xt, queries):
        A = _i64(A)
        Xt = _i64(Xt)
        if A.ndim == 2:
            A = A[None]
            Xt = Xt[None]
        q = np.ascontiguousarray(queries, dtype=np.int32).reshape(-1, 4)
        out = np.zeros((len(q), 6), np.int64)
        self._check(self._lib.xtx_probe_moves(self._h, _p64(A), _p64(Xt), A.shape[1], A.shape[0],
                                              q.ctypes.data_as(N.I32P), len(q), _p64(out)))
        return out

    def probe_rng(self, ctr_key, xs):
        ck = np.ascontiguousarray(ctr_key, dtype=np.uint32).reshape(-1, 6)
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        out = np.zeros((len(ck), 4), np.uint32)
        ex = np.zeros(len(xs), np.float64)
        self._check(self._lib.xtx_probe_rng(
            self._h, ck.ctypes.data_as(C.POINTER(C.c_uint32)), len(ck),
            out.ctypes.data_as(C.POINTER(C.c_uint32)),
            xs.ctypes.data_as(C.POINTER(C.c_double)), len(xs),
            ex.ctypes.data_as(C.POINTER(C.c_double))))
        return out, ex


def verify(A0, X, A):
    """Exact host verification (xtx_verify): returns (status, det_sign)."""
    A0 = _i64(A0)
    X = _i64(X)
    A = _i64(A)
    rep = N.VerifyReport()
    rc = N.lib().xtx_verify(_p64(A0), _p64(X), _p64(A), A0.shape[0], C.byref(rep))
    if rc != N.XTX_OK:
        raise N.XtxError(rc, "xtx_verify")
    return rep.status, rep.det_sign
