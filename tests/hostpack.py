"""Pack oracle blocks into the product's arena layout (test helper)."""
import ctypes as C

import numpy as np

from hofstadter_hubbard_tsc_amd import _lib as L

KIND_NAMES = ["cdag", "S", "n"]


def desc(block):
    N = np.array([q[0] for q in block.sectors], np.int32)
    S2 = np.array([q[1] for q in block.sectors], np.int32)
    d = np.array(block.dims, np.int32)
    sites = np.array(sorted(block.ops.keys()), np.int32)
    return N, S2, d, sites


def plan_matvec(model, sysb, envb, N):
    lib = L.load()
    mp = L.model_params(model)
    sN, sS, sd, ss = desc(sysb)
    eN, eS, ed, es = desc(envb)
    p = C.c_void_p()
    rc = lib.hh_plan_matvec(C.byref(mp), sysb.length, len(sN), L.ip(sN), L.ip(sS), L.ip(sd), len(ss), L.ip(ss),
                            envb.length, len(eN), L.ip(eN), L.ip(eS), L.ip(ed), len(es), L.ip(es), N, C.byref(p))
    assert rc == 0, lib.hh_last_error(None)
    return p


def layout(plan, which, kind):
    lib = L.load()
    n = L.i32()
    lib.hh_plan_layout(plan, which, kind, C.byref(n), None, None, None)
    bra, ket = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
    off = np.zeros(n.value, np.int64)
    lib.hh_plan_layout(plan, which, kind, C.byref(n), L.ip(bra), L.ip(ket), L.ip(off))
    return bra, ket, off


def block_info(plan, which):
    lib = L.load()
    a, ns, nst = L.i64(), L.i32(), L.i32()
    lib.hh_plan_block(plan, which, C.byref(a), C.byref(ns), None, None, None, None, C.byref(nst), None)
    N, S2, d = (np.zeros(ns.value, np.int32) for _ in range(3))
    h = np.zeros(ns.value, np.int64)
    sites = np.zeros(nst.value, np.int32)
    lib.hh_plan_block(plan, which, C.byref(a), C.byref(ns), L.ip(N), L.ip(S2), L.ip(d), L.ip(h), C.byref(nst),
                      L.ip(sites))
    return a.value, N, S2, d, h, sites


def op_offset(plan, which, si, kind):
    o = L.i64()
    rc = L.load().hh_plan_op_offset(plan, which, si, kind, C.byref(o))
    return o.value if rc == 0 else -1


def pack(plan, which, block, nkinds):
    """Arena of `block` (an oracle Block) in the plan's layout."""
    size, N, S2, d, h, sites = block_info(plan, which)
    arena = np.zeros(max(size, 1), np.complex128)
    for i, dim in enumerate(d):
        arena[h[i]:h[i] + dim * dim] = block.H.blk[(i, i)].reshape(-1, order="F")
    for si, k in enumerate(sites):
        for kind in range(nkinds):
            base = op_offset(plan, which, si, kind)
            bra, ket, off = layout(plan, which, kind)
            op = block.ops[int(k)][KIND_NAMES[kind]]
            for b_, k_, o in zip(bra, ket, off):
                m = op.blk.get((int(b_), int(k_)))
                if m is not None:
                    arena[base + o: base + o + m.size] = m.reshape(-1, order="F")
    return arena


def unpack_block(plan, which, arena, nkinds):
    """(sectors, H dict, ops dict site->kind->{(bra,ket): mat}) from an arena."""
    size, N, S2, d, h, sites = block_info(plan, which)
    H = {i: arena[h[i]:h[i] + dim * dim].reshape((dim, dim), order="F") for i, dim in enumerate(d)}
    ops = {}
    for si, k in enumerate(sites):
        ops[int(k)] = {}
        for kind in range(nkinds):
            base = op_offset(plan, which, si, kind)
            bra, ket, off = layout(plan, which, kind)
            blk = {}
            for b_, k_, o in zip(bra, ket, off):
                r, c = d[b_], d[k_]
                blk[(int(b_), int(k_))] = arena[base + o: base + o + r * c].reshape((r, c), order="F")
            ops[int(k)][KIND_NAMES[kind]] = blk
    return list(zip(N.tolist(), S2.tolist(), d.tolist())), H, ops
