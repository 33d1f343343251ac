"""The C-ABI library builds, loads without a GPU, and exports every function
that include/swiftest_hip.h declares; the product path refuses to run
without its native library."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "swiftest_hip.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(swg_\w+)\s*\(", src, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ("swg_init", "swg_kick_getacch_int_all_tri_pl", "swg_kick_getacch_int_all_tp", "swg_drift_all",
              "swg_encounter_check_all_plpl", "swg_system_create", "swg_system_run"):
        assert n in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "swiftest_amd", "libswiftest_hip.so")
    if not os.path.exists(path):
        from swiftest_amd import build
        build.build()
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_init_without_gpu_fails_cleanly():
    import swiftest_amd.core as core
    lib = core.library()
    h = ctypes.c_void_p()
    rc = lib.swg_init(999, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert lib.swg_last_error()


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    import swiftest_amd.core as core
    monkeypatch.setattr(core, "_lib", None)
    monkeypatch.setattr(core, "LIBPATH", "/nonexistent/libswiftest_hip.so")
    with pytest.raises(core.SwiftestHipError):
        core.library()
