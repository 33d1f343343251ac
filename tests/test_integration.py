"""The Fortran interface module of the integration notes (the `module
swiftest_hip` block of the repository's integration document) compiles with flang,
links against the library, and its bind(c) interfaces reach the C ABI: an
init on a device that does not exist returns an error code and a message,
and the derived types have the C structs' sizes."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"

PROGRAM = """
program check_binding
   use swiftest_hip
   use, intrinsic :: iso_c_binding
   implicit none
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   character(kind=c_char), pointer :: msg(:)
   type(swg_params) :: prm
   type(swg_stats) :: st
   ctx = c_null_ptr
   rc = swg_init(999_c_int, ctx)
   call c_f_pointer(swg_last_error(), msg, [200])
   print '(a,i0)', 'rc=', rc
   print '(a,l1)', 'ctx_null=', .not. c_associated(ctx)
   print '(a,a1)', 'msg0=', msg(1)
   print '(a,i0)', 'sizeof_params=', c_sizeof(prm)
   print '(a,i0)', 'sizeof_stats=', c_sizeof(st)
end program check_binding
"""


def interface_module():
    notes = os.path.join(ROOT, "INTEGRATION.md")
    text = open(notes).read()
    blocks = re.findall(r"```fortran\n(.*?)```", text, re.S)
    mod = [b for b in blocks if b.lstrip().startswith("module swiftest_hip")]
    assert mod, f"no interface module in {notes}"
    return mod[0]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_interface_module_links_and_calls(tmp_path):
    lib = os.path.join(ROOT, "swiftest_amd", "libswiftest_hip.so")
    if not os.path.exists(lib):
        from swiftest_amd import build
        build.build()
    (tmp_path / "swiftest_hip_mod.f90").write_text(interface_module())
    (tmp_path / "main.f90").write_text(PROGRAM)
    exe = tmp_path / "check_binding"
    libdir = os.path.dirname(lib)
    cmd = [FLANG, "-o", str(exe), str(tmp_path / "swiftest_hip_mod.f90"), str(tmp_path / "main.f90"),
           f"-L{libdir}", "-lswiftest_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = dict(line.strip().split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    assert int(out["rc"]) < 0
    assert out["ctx_null"] == "T"
    assert out["msg0"].strip() != ""
    # struct sizes must equal the C header's (swg_params 8+5*4+3*4 = 40; swg_stats 56)
    import ctypes
    import swiftest_amd.core as core
    assert int(out["sizeof_params"]) == ctypes.sizeof(core.Params)
    assert int(out["sizeof_stats"]) == ctypes.sizeof(core.Stats)
