"""Host builds of the device arithmetic helpers, checked against IEEE results.

The constant-divisor quotient cdiv (swg_math.h) replaces x / c in the Stumpff
series (swiftest_drift.f90:536-580), kepmd (:234-276) and kepu_new (:435-483).
Its host instantiation is compiled with hipcc and compared bit for bit with
the IEEE quotient over random operands spanning the normal range; the device
instantiation is covered by the bit-exact drift parity tests.  The same
program compares the compile-time table of RSHELL powers (swg_level.h; the
shell factors of symba_kick.f90:150-200 and the encounter radii of
symba_util.f90:245-267) with the powers evaluated at run time, and checks
pow_third, the correctly rounded x**THIRD of the cubic first guess
(swiftest_drift.f90:486-533), against 80-bit powl and on its exact cubes and
out-of-range fallbacks.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not installed")
def test_device_math_helpers(tmp_path):
    exe = str(tmp_path / "math_check")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-O2", "-std=c++17", "-ffp-contract=off",
           "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(HERE, "native", "math_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok "), out.stdout
