"""swiftest_amd -- MI355X implementation of Swiftest's per-step hot path.

The compute path is libswiftest_hip.so (hand-written HIP kernels for gfx950
behind the C ABI declared in include/swiftest_hip.h).  This package is the
Python host mirror of the reference's operator API over that ABI, plus the
offline initial-condition generator used by the benchmarks and tests.
"""
from .core import (HELIO, PAIR_EXACT, PAIR_FAST, SYMBA, WHM, GpuSystem, SwiftestHipError, context,
                   encounter_check_all_plpl, encounter_check_all_pltp, library, set_pair_mode, swiftest_drift_all,
                   swiftest_kick_getacch_int_all_flat, swiftest_kick_getacch_int_all_tp,
                   swiftest_kick_getacch_int_all_tri)

__all__ = ["WHM", "HELIO", "SYMBA", "PAIR_EXACT", "PAIR_FAST", "set_pair_mode", "GpuSystem", "SwiftestHipError", "context", "library",
           "swiftest_kick_getacch_int_all_tri", "swiftest_kick_getacch_int_all_flat",
           "swiftest_kick_getacch_int_all_tp", "swiftest_drift_all", "encounter_check_all_plpl",
           "encounter_check_all_pltp"]
