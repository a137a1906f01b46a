"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/hhdmrg.h declares; the Python binding covers all of them; host-only
error paths return status codes instead of aborting."""
import ctypes as C
import os
import re

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


def test_ctx_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        return
    lib = L.load()
    p = C.c_void_p()
    rc = lib.hh_ctx_create(0, C.byref(p))
    assert rc != 0
    assert lib.hh_last_error(None)
