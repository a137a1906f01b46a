"""numpy executor of the engine's task lists (test helper: checks the host
task compilers on the CPU; the product runs these lists on the GPU)."""
import ctypes as C

import numpy as np

from hofstadter_hubbard_tsc_amd import _lib as L


def fetch(plan_ptr, stage, fn="hh_plan_tasks"):
    lib = L.load()
    tp, np_, tm, nm, sc = C.c_void_p(), L.i32(), C.c_void_p(), L.i32(), L.i64()
    rc = getattr(lib, fn)(plan_ptr, stage, C.byref(tp), C.byref(np_), C.byref(tm), C.byref(nm), C.byref(sc))
    assert rc == 0
    tiles = np.frombuffer((C.c_char * (np_.value * L.TILE_DTYPE.itemsize)).from_address(tp.value or 0),
                          dtype=L.TILE_DTYPE).copy() if np_.value else np.zeros(0, L.TILE_DTYPE)
    terms = np.frombuffer((C.c_char * (nm.value * L.TERM_DTYPE.itemsize)).from_address(tm.value or 0),
                          dtype=L.TERM_DTYPE).copy() if nm.value else np.zeros(0, L.TERM_DTYPE)
    return tiles, terms, sc.value


def _view(buf, off, ld, rows, cols):
    idx = off + np.arange(rows)[:, None] + ld * np.arange(cols)[None, :]
    return buf[idx]


def _op(buf, off, ld, op, rows, cols):
    """op(A) as a rows x cols matrix (op 0 N, 1 T, 2 C, 3 conj)."""
    if op in (0, 3):
        m = _view(buf, off, ld, rows, cols)
    else:
        m = _view(buf, off, ld, cols, rows).T
    if op in (2, 3):
        m = np.conj(m)
    return m


def run(tiles, terms, bufs):
    for t in tiles:
        M, N = int(t["M"]), int(t["N"])
        acc = np.zeros((M, N), np.complex128)
        for k in range(int(t["t0"]), int(t["t1"])):
            tm = terms[k]
            c = complex(tm["cr"], tm["ci"])
            K = int(tm["K"])
            A = bufs[int(tm["a_buf"])]
            if K == 0:
                acc += c * _op(A, int(tm["a_off"]), int(tm["a_ld"]), int(tm["a_op"]), M, N)
            else:
                B = bufs[int(tm["b_buf"])]
                a = _op(A, int(tm["a_off"]), int(tm["a_ld"]), int(tm["a_op"]), M, K)
                b = _op(B, int(tm["b_off"]), int(tm["b_ld"]), int(tm["b_op"]), K, N)
                acc += c * (a @ b)
        out = bufs[int(t["c_buf"])]
        idx = int(t["c_off"]) + np.arange(M)[:, None] + int(t["c_ld"]) * np.arange(N)[None, :]
        if int(t["accumulate"]):
            out[idx] += acc
        else:
            out[idx] = acc


def _sview(buf, off, ld, rows, cols, writeable=False):
    """rows x cols column-major view of buf at element off with leading
    dimension ld, without a copy (the task compilers keep every operand in
    bounds)."""
    it = buf.itemsize
    return np.lib.stride_tricks.as_strided(buf[off:], shape=(rows, cols), strides=(it, it * ld),
                                           writeable=writeable)


def _sop(buf, off, ld, op, rows, cols):
    if op in (0, 3):
        m = _sview(buf, off, ld, rows, cols)
    else:
        m = _sview(buf, off, ld, cols, rows).T
    if op in (2, 3):
        m = np.conj(m)
    return m


def run_strided(tiles, terms, bufs):
    """run() on strided views instead of gathered copies: the same sector
    GEMMs in the same order, fast enough for full-size task lists."""
    tl = tiles.tolist()
    tm = terms.tolist()
    names = terms.dtype.names
    ix = {n: i for i, n in enumerate(names)}
    tn = {n: i for i, n in enumerate(tiles.dtype.names)}
    for t in tl:
        M, N = t[tn["M"]], t[tn["N"]]
        acc = np.zeros((M, N), np.complex128)
        for k in range(t[tn["t0"]], t[tn["t1"]]):
            x = tm[k]
            c = complex(x[ix["cr"]], x[ix["ci"]])
            K = x[ix["K"]]
            A = bufs[x[ix["a_buf"]]]
            if K == 0:
                acc += c * _sop(A, x[ix["a_off"]], x[ix["a_ld"]], x[ix["a_op"]], M, N)
            else:
                B = bufs[x[ix["b_buf"]]]
                a = _sop(A, x[ix["a_off"]], x[ix["a_ld"]], x[ix["a_op"]], M, K)
                b = _sop(B, x[ix["b_off"]], x[ix["b_ld"]], x[ix["b_op"]], K, N)
                acc += c * (a @ b)
        out = _sview(bufs[t[tn["c_buf"]]], t[tn["c_off"]], t[tn["c_ld"]], M, N, writeable=True)
        if t[tn["accumulate"]]:
            out += acc
        else:
            out[...] = acc
