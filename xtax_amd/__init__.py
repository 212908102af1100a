"""xtax_amd: MI355X (gfx950) engine for the xtax L1 congruence annealer.

The hot path is HIP (xtax_amd/csrc/anneal.hip) behind the C ABI of
include/xtx.h; this package is the Python host binding used by tests and the
benchmark, and xtax_amd/xtax is the C++20 command-line drop-in.
"""
from .engine import DualParams, DualResult, Engine, EngineParams, SolveResult, verify  # noqa: F401
from . import workloads  # noqa: F401

__all__ = ["DualParams", "DualResult", "Engine", "EngineParams", "SolveResult", "verify",
           "workloads"]
