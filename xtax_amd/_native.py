"""ctypes binding of include/xtx.h (libxtx.so, built in-tree by xtax_amd/Makefile).

There is no fallback: if the gfx950 library is missing, importing the engine
raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxtx.so")

I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)

XTX_OK = 0
ERRORS = {-1: "XTX_E_ARG", -2: "XTX_E_BOUND", -3: "XTX_E_DEVICE", -4: "XTX_E_NOMEM",
          -5: "XTX_E_STATE", -6: "XTX_E_FULL", -7: "XTX_E_VERIFY"}


class Params(C.Structure):
    _fields_ = [
        ("t_init", C.c_double), ("t_min", C.c_double), ("cooling", C.c_double),
        ("reheat", C.c_double), ("reseed_factor", C.c_double),
        ("greedy_fraction", C.c_double), ("target_fraction", C.c_double),
        ("target_accept", C.c_double), ("worker_diversity", C.c_double),
        ("sweep_fraction", C.c_double), ("max_seconds", C.c_double),
        ("save_interval", C.c_double),
        ("moves_per_cool", C.c_int32), ("stuck_threshold", C.c_int32),
        ("target_samples", C.c_int32), ("tempering", C.c_int32),
        ("adaptive_cooling", C.c_int32), ("exchange_interval", C.c_int32),
        ("seed", C.c_uint64),
        ("replicas", C.c_int32), ("move_log", C.c_int32), ("move_log_cap", C.c_int64),
        ("max_moves", C.c_uint64), ("max_epochs", C.c_int32), ("verify", C.c_int32),
        ("waves_per_block", C.c_int32), ("autotune", C.c_int32), ("lookahead", C.c_int32),
        ("persistent", C.c_int32),
        ("on_best", C.c_void_p), ("on_best_user", C.c_void_p),
    ]


# void (*on_best)(void* user, double seconds, int64_t score, double dscore, int32_t replica)
ON_BEST = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_int64, C.c_double, C.c_int32)


class Limits(C.Structure):
    _fields_ = [("compute_units", C.c_int32), ("lds_bytes_per_cu", C.c_int32),
                ("lds_bytes_per_block_max", C.c_int32), ("wave_size", C.c_int32),
                ("max_n", C.c_int32), ("max_n_lds", C.c_int32), ("hbm_bytes", C.c_int64),
                ("arch", C.c_char * 64)]


class Stats(C.Structure):
    _fields_ = [
        ("moves", C.c_int64), ("accepts", C.c_int64), ("greedy_calls", C.c_int64),
        ("infeasible", C.c_int64), ("sweeps", C.c_int64), ("exchanges_tried", C.c_int64),
        ("exchanges_accepted", C.c_int64), ("solved", C.c_int32),
        ("solve_replica", C.c_int32), ("solve_move", C.c_int64), ("best_score", C.c_int64),
        ("best_replica", C.c_int32), ("epochs", C.c_int32), ("stop_reason", C.c_int32),
        ("verify_status", C.c_int32), ("seconds", C.c_double), ("kernel_ms", C.c_double),
    ]


class Instance(C.Structure):
    _fields_ = [("A0", I64P), ("X0", I64P), ("n", C.c_int32), ("replicas", C.c_int32),
                ("seed", C.c_uint64), ("active", I32P), ("n_active", C.c_int32),
                ("reserved", C.c_int32), ("A_out", I64P), ("X_out", I64P)]


class Result(C.Structure):
    _fields_ = [("solved", C.c_int32), ("verify_status", C.c_int32),
                ("solve_move", C.c_int64), ("best_score", C.c_int64), ("moves", C.c_int64),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("retire_s", C.c_double),
                ("tag", C.c_uint64)]


class BatchStats(C.Structure):
    _fields_ = [("moves", C.c_int64), ("accepts", C.c_int64), ("instances", C.c_int32),
                ("solved", C.c_int32), ("verified", C.c_int32), ("epochs", C.c_int32),
                ("seconds", C.c_double), ("kernel_ms", C.c_double), ("buckets", C.c_int32),
                ("graph_epochs", C.c_int32), ("upload_seconds", C.c_double),
                ("lane_instances", C.c_int32), ("persistent", C.c_int32)]


class ServeConfig(C.Structure):
    _fields_ = [("lds_budget_bytes", C.c_int64), ("hbm_budget_bytes", C.c_int64),
                ("max_resident", C.c_int32), ("queue_capacity", C.c_int32)]


class VerifyReport(C.Structure):
    _fields_ = [("status", C.c_int32), ("primes_used", C.c_int32), ("det_sign", C.c_int32),
                ("exact_int128", C.c_int32)]


class Launch(C.Structure):
    _fields_ = [("order", C.c_int32), ("lanes", C.c_int32), ("cols", C.c_int32),
                ("triangular", C.c_int32), ("waves_per_block", C.c_int32),
                ("tuned", C.c_int32), ("lds_bytes", C.c_int32), ("lookahead", C.c_int32),
                ("x_in_lds", C.c_int32), ("reserved", C.c_int32), ("moves_per_sec", C.c_double)]


class DualParams(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("dual_tol", C.c_double), ("lambda_ramp", C.c_double),
                ("dual_refresh", C.c_int32), ("dual_check", C.c_int32)]


class DualStats(C.Structure):
    _fields_ = [
        ("moves", C.c_int64), ("accepts", C.c_int64), ("greedy_calls", C.c_int64),
        ("infeasible", C.c_int64), ("sweeps", C.c_int64), ("exchanges_tried", C.c_int64),
        ("exchanges_accepted", C.c_int64), ("refreshes", C.c_int64), ("checks", C.c_int64),
        ("recomputes", C.c_int64), ("solved", C.c_int32), ("solve_replica", C.c_int32),
        ("solve_move", C.c_int64), ("best_score", C.c_double), ("start_score", C.c_double),
        ("start_primal", C.c_double), ("start_dual", C.c_double),
        ("c", C.c_double), ("dual_residual", C.c_double), ("best_replica", C.c_int32),
        ("epochs", C.c_int32), ("stop_reason", C.c_int32), ("verify_status", C.c_int32),
        ("seconds", C.c_double), ("kernel_ms", C.c_double),
    ]


class LogEntry(C.Structure):
    _fields_ = [("move", C.c_int64), ("i", C.c_int16), ("j", C.c_int16), ("s", C.c_int16),
                ("flags", C.c_uint16)]


# Every symbol include/xtx.h declares, with (restype, argtypes).
SIGNATURES = {
    "xtx_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "xtx_close": (C.c_int, [C.c_void_p]),
    "xtx_last_error": (C.c_char_p, [C.c_void_p]),
    "xtx_query_limits": (C.c_int, [C.c_void_p, C.POINTER(Limits)]),
    "xtx_params_default": (None, [C.POINTER(Params)]),
    "xtx_request_stop": (None, [C.c_void_p]),
    "xtx_clear_stop": (None, [C.c_void_p]),
    "xtx_launch_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Launch)]),
    "xtx_launch_set": (C.c_int, [C.c_void_p, C.POINTER(Launch)]),
    "xtx_autotune": (C.c_int, [C.c_void_p, C.POINTER(Instance), C.c_int32, C.POINTER(Params),
                               C.c_int32, C.POINTER(Launch), C.c_int32, I32P]),
    "xtx_solve_plain": (C.c_int, [C.c_void_p, I64P, I64P, C.c_int32, C.POINTER(Params), I32P,
                                  C.c_int32, I64P, I64P, C.POINTER(Stats)]),
    "xtx_solve_plain_noalloc": (C.c_int, [C.c_void_p, I64P, I64P, C.c_int32,
                                          C.POINTER(Params), I32P, C.c_int32, I64P, I64P,
                                          C.POINTER(Stats)]),
    "xtx_last_replicas": (C.c_int, [C.c_void_p, I64P, I64P, I64P]),
    "xtx_best_so_far": (C.c_int, [C.c_void_p, I64P, C.POINTER(C.c_double), I64P]),
    "xtx_movelog_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.POINTER(LogEntry)),
                                  I64P, I32P]),
    "xtx_movelog_free": (C.c_int, [C.c_void_p]),
    "xtx_batch_solve": (C.c_int, [C.c_void_p, C.POINTER(Instance), C.c_int32,
                                  C.POINTER(Params), C.POINTER(Result),
                                  C.POINTER(BatchStats)]),
    "xtx_serve_open": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(ServeConfig)]),
    "xtx_serve_submit": (C.c_int, [C.c_void_p, C.POINTER(Instance), C.c_uint64]),
    "xtx_serve_step": (C.c_int, [C.c_void_p, I32P, I32P]),
    "xtx_serve_poll": (C.c_int, [C.c_void_p, C.POINTER(Result), C.c_int32, I32P]),
    "xtx_serve_retire": (C.c_int, [C.c_void_p, C.c_uint64]),
    "xtx_serve_close": (C.c_int, [C.c_void_p]),
    "xtx_verify": (C.c_int, [I64P, I64P, I64P, C.c_int32, C.POINTER(VerifyReport)]),
    "xtx_reorder": (C.c_int, [I64P, I64P, C.c_int32, C.c_int32, I32P]),
    "xtx_is_unimodular": (C.c_int, [I64P, C.c_int32, I32P]),
    "xtx_deflate_pivot": (C.c_int, [I64P, I64P, C.c_int32, C.c_int32, I32P, C.c_int32, I32P]),
    "xtx_dual_params_default": (None, [C.POINTER(DualParams)]),
    "xtx_dual_solve": (C.c_int, [C.c_void_p, I64P, I64P, C.c_int32, C.POINTER(Params),
                                 C.POINTER(DualParams), I32P, C.c_int32, I64P,
                                 C.POINTER(C.c_double), I64P, C.POINTER(DualStats)]),
    "xtx_dual_last_replicas": (C.c_int, [C.c_void_p, I64P, C.POINTER(C.c_double), I64P,
                                         C.POINTER(C.c_double), I64P]),
    "xtx_dual_probe": (C.c_int, [C.c_void_p, I64P, C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                 C.c_double, I32P, C.c_int32, I64P, C.POINTER(C.c_double), I64P,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "xtx_probe_moves": (C.c_int, [C.c_void_p, I64P, I64P, C.c_int32, C.c_int32, I32P,
                                  C.c_int32, I64P]),
    "xtx_probe_rng": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_int32,
                                C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.c_int32,
                                C.POINTER(C.c_double)]),
}

_LIB = None


def lib():
    """The loaded libxtx.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C xtax_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class XtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
