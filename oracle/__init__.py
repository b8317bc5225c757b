"""CPU oracle for the RheoCell per-step hot path -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg, and there only as the checker.  The product package
``rheocell_amd`` never imports it.
"""
from .pyoracle import OracleSim, OrcParams, params_from_keys, lib_path  # noqa: F401
