"""The allocation-free engine entry point (hh_run_tasks_noalloc) against
hh_run_tasks through the C ABI: identical output on two calls in a row, the
second with a smaller task list (the fixed scratch keeps the first call's
contents, which must not leak into the second), both against the numpy
executor of the task lists; a task list too large for the scratch is refused
with status 1 and nothing launched."""
import ctypes as C

import numpy as np
import pytest

from . import npexec
from .gpuhelp import dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a device"
    from hofstadter_hubbard_tsc_amd import Context
    c = Context(0)
    yield c
    c.close()


def _rand(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _task_list(rng, shapes, nterm_ops):
    """Tiles over buffer slot 1 (output), terms reading slots 2 (A) and 3 (B):
    every operand op, ragged shapes, K = 0 elementwise terms, the accumulate
    flag, and term lists long enough to be split over several waves (their
    partial sums go through the scratch)."""
    from hofstadter_hubbard_tsc_amd import _lib as L
    tiles, terms = [], []
    off = 0
    for (M, N, K) in shapes:
        for acc in (0, 1):
            t0 = len(terms)
            for rep in range(nterm_ops):
                for aop in range(4):
                    for bop in ((0, 1, 2, 3) if K else (0,)):
                        lda = (M if aop in (0, 3) else K) + 3 if K else M + 2
                        ldb = (K if bop in (0, 3) else N) + 1
                        terms.append((rng.normal(), rng.normal(), int(rng.integers(0, 2000)),
                                      int(rng.integers(0, 2000)), lda, ldb, K, 2, 3,
                                      aop if K else (aop if aop in (0, 3) else 0), bop))
            tiles.append((off, M + 5, M, N, t0, len(terms), 1, acc))
            off += (M + 5) * N + 7
    return np.array(tiles, dtype=L.TILE_DTYPE), np.array(terms, dtype=L.TERM_DTYPE), off


def _run(ctx, fn, T, R, A, B, out0, precision=0):
    from hofstadter_hubbard_tsc_amd import _lib as L
    dA, dB, dO = dev(A), dev(B), dev(out0)
    bufs = (C.c_void_p * 8)(None, dO.data_ptr(), dA.data_ptr(), dB.data_ptr(), None, None, None, None)
    rc = getattr(L.load(), fn)(ctx.ptr, T.ctypes.data_as(C.c_void_p), len(T), R.ctypes.data_as(C.c_void_p), len(R),
                               bufs, precision)
    return rc, host(dO)


def test_noalloc_matches_run_tasks_on_two_calls(ctx):
    rng = np.random.default_rng(23)
    A = _rand(rng, 400000)
    B = _rand(rng, 400000)
    big = [(1, 1, 1), (16, 16, 4), (70, 45, 37), (64, 64, 64), (129, 3, 17), (5, 130, 9), (33, 33, 0),
           (200, 150, 48)]
    small = [(16, 16, 4), (37, 21, 5), (33, 33, 0)]
    for call, (shapes, reps) in enumerate(((big, 3), (small, 1))):
        T, R, size = _task_list(rng, shapes, reps)
        out0 = _rand(rng, size + 100)
        ref = out0.copy()
        npexec.run(T, R, [None, ref, A, B, None, None, None, None])
        rc_a, got_a = _run(ctx, "hh_run_tasks", T, R, A, B, out0)
        rc_b, got_b = _run(ctx, "hh_run_tasks_noalloc", T, R, A, B, out0)
        print("call %d: %d tiles, %d terms" % (call + 1, len(T), len(R)), flush=True)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        scale = np.abs(ref).max()
        # same task list, same kernels, same work split: the same bits
        assert np.array_equal(got_a, got_b), np.abs(got_a - got_b).max()
        assert np.abs(got_b - ref).max() / scale < 1e-13


def test_noalloc_refuses_a_task_list_beyond_the_scratch(ctx):
    from hofstadter_hubbard_tsc_amd import _lib as L
    lib = L.load()
    # 2^21 one-element tiles with one elementwise term each: 2^21 * (40 + 48)
    # bytes of tables, beyond the 64 MiB scratch
    n = 1 << 21
    T = np.zeros(n, dtype=L.TILE_DTYPE)
    T["c_off"] = np.arange(n)
    T["c_ld"] = 1
    T["M"] = 1
    T["N"] = 1
    T["t0"] = np.arange(n)
    T["t1"] = np.arange(n) + 1
    T["c_buf"] = 1
    R = np.zeros(n, dtype=L.TERM_DTYPE)
    R["cr"] = 1.0
    R["a_off"] = np.arange(n)
    R["a_ld"] = 1
    R["a_buf"] = 2
    x = dev(np.zeros(n, np.complex128))
    y = dev(np.zeros(n, np.complex128))
    bufs = (C.c_void_p * 8)(None, y.data_ptr(), x.data_ptr(), None, None, None, None, None)
    rc = lib.hh_run_tasks_noalloc(ctx.ptr, T.ctypes.data_as(C.c_void_p), n, R.ctypes.data_as(C.c_void_p), n, bufs, 0)
    assert rc == 1
    assert b"scratch" in lib.hh_last_error(ctx.ptr)
