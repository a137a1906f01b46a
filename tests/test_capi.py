"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/hhdmrg.h declares; the Python binding covers all of them; host-only
error paths return status codes instead of aborting."""
import ctypes as C
import os
import re

import pytest

from hofstadter_hubbard_tsc_amd import _lib as L

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "hhdmrg.h")


def _declared():
    src = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hh_\w+)\s*\(", src, flags=re.M)))


def test_header_symbols_exported():
    lib = L.load()
    names = _declared()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(_declared()) <= bound


def test_plan_rejects_nothing_silently():
    # a plan with a bad model kind still returns a status, never aborts
    lib = L.load()
    mp = L.ModelParams(0, 2, 2, 0, 0, 2, 1.0, 8.0, 0.0, 0.0)
    one = (C.c_int32 * 1)(1)
    z = (C.c_int32 * 1)(0)
    p = C.c_void_p()
    rc = lib.hh_plan_matvec(C.byref(mp), 1, 1, z, z, one, 1, one, 1, 1, z, z, one, 1, one, 0, C.byref(p))
    assert rc == 0
    lib.hh_plan_destroy(p)


@pytest.mark.parametrize("fields,msg", [
    (dict(lattice=0, gauge=1, theta=1.0), b"triangular"),           # gauge 1 on the square lattice
    (dict(lattice=1, gauge=1, theta=1.0, flux=0.1), b"flux"),        # gauge 1 with a Landau flux
    (dict(lattice=1, gauge=2), b"gauge"),
    (dict(lattice=1, gauge=0, reserved=3), b"reserved"),
])
def test_plan_refuses_bad_gauges(fields, msg):
    lib = L.load()
    mp = L.ModelParams(0, 2, 2, 0, 0, 2, 1.0, 8.0, 0.0, 0.0)
    for k, v in fields.items():
        setattr(mp, k, v)
    one = (C.c_int32 * 1)(1)
    z = (C.c_int32 * 1)(0)
    p = C.c_void_p()
    rc = lib.hh_plan_matvec(C.byref(mp), 1, 1, z, z, one, 1, one, 1, 1, z, z, one, 1, one, 0, C.byref(p))
    assert rc != 0
    assert msg in lib.hh_last_error(None)


def test_ctx_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        return
    lib = L.load()
    p = C.c_void_p()
    rc = lib.hh_ctx_create(0, C.byref(p))
    assert rc != 0
    assert lib.hh_last_error(None)


def test_header_compiles_as_c99_and_cxx(tmp_path):
    """include/hhdmrg.h is a plain C header: it compiles as pedantic C99 (no
    duplicate typedefs, no C++ in the declarations) and as C++11."""
    import shutil
    import subprocess
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src_c = tmp_path / "h.c"
    src_c.write_text('#include "hhdmrg.h"\nint main(void) { return 0; }\n')
    src_cc = tmp_path / "h.cc"
    src_cc.write_text('#include "hhdmrg.h"\nint main() { return 0; }\n')
    for cc, std, src in (("gcc", "-std=c99", src_c), ("g++", "-std=c++11", src_cc)):
        if shutil.which(cc) is None:
            pytest.skip("%s not available" % cc)
        r = subprocess.run([cc, std, "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", inc, "-c", str(src), "-o",
                            str(tmp_path / "h.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
