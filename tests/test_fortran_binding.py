"""The Fortran iso_c_binding module of INTEGRATION.md, compiled with the
image's flang: a small Fortran caller links against libhhdmrg.so, gets the
error status of hh_ctx_create on a machine without a device together with its
message through hh_last_error, and sees the same sizes of the bind(C) types as
the C header (the ctypes mirror of the same structs)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from hofstadter_hubbard_tsc_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

CALLER = """program caller
  use hhdmrg
  use iso_c_binding
  implicit none
  type(c_ptr) :: ctx, msg
  integer(c_int) :: rc
  type(hh_step_record) :: r
  type(hh_schedule) :: s
  type(hh_model_params) :: m
  character(kind=c_char), pointer :: text(:)
  integer :: i
  rc = hh_ctx_create(0, ctx)
  print '(A,I0)', 'rc ', rc
  msg = hh_last_error(c_null_ptr)
  call c_f_pointer(msg, text, [200])
  i = 1
  do while (i <= 200)
     if (text(i) == c_null_char) exit
     i = i + 1
  end do
  print '(A,I0)', 'msglen ', i - 1
  print '(A,I0)', 'record ', storage_size(r) / 8
  print '(A,I0)', 'schedule ', storage_size(s) / 8
  print '(A,I0)', 'model ', storage_size(m) / 8
end program caller
"""


def test_fortran_module_compiles_links_and_matches_the_header(tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler in this image")
    import torch
    if torch.cuda.is_available():
        pytest.skip("checks the no-device error path")
    L.load()
    src = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    mod = re.search(r"```fortran\n(module hhdmrg.*?end module hhdmrg)\n```", src, flags=re.S).group(1)
    (tmp_path / "hhdmrg_mod.f90").write_text(mod + "\n")
    (tmp_path / "caller.f90").write_text(CALLER)
    libdir = os.path.dirname(L.LIB_PATH)
    exe = tmp_path / "caller"
    r = subprocess.run([FLANG, "-o", str(exe), str(tmp_path / "hhdmrg_mod.f90"), str(tmp_path / "caller.f90"),
                        "-L" + libdir, "-lhhdmrg", "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    out = dict(line.split() for line in run.stdout.strip().splitlines())
    assert int(out["rc"]) != 0
    assert int(out["msglen"]) > 0
    assert int(out["record"]) == C.sizeof(L.StepRecord)
    assert int(out["schedule"]) == C.sizeof(L.Schedule)
    assert int(out["model"]) == C.sizeof(L.ModelParams)


def test_c_struct_sizes_match_the_python_binding(tmp_path):
    """sizeof of every struct of include/hhdmrg.h, compiled with gcc, equals
    the ctypes mirror's size and field offsets."""
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"hh_step_record": L.StepRecord, "hh_schedule": L.Schedule, "hh_model_params": L.ModelParams,
               "hh_stats": L.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hhdmrg.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append('  return 0; }')
    (tmp_path / "s.c").write_text("\n".join(lines) + "\n")
    exe = tmp_path / "s"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(tmp_path / "s.c")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out["%s.%s" % (cname, fname)]) == getattr(py, fname).offset, (cname, fname)
