"""The enlarged-basis and super-basis order of the product's host logic
against a fixture derived by hand from the reference's Get_basis,
QuickSort_Integer, Get_Unique_QN and Get_super_basis (code/update.f90:11-273):
sector order 2S descending then N descending, subsectors block-major with
running spos, super sectors sys-major.  Also checked for the oracle, and a
planted swap of the sort keys must not match the fixture."""
import ctypes as C
import json
import os

import numpy as np

from hofstadter_hubbard_tsc_amd import _lib as L
from oracle import dmrg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "basis_order_hubbard.json")


def _fixture():
    return json.load(open(GOLD))


def _product(fx):
    lib = L.load()
    mp = L.ModelParams(0, fx["model"]["Lx"], fx["model"]["Ly"], 0, 1, fx["model"]["n_elec"], 1.0, 8.0, 0.0, 0.0)
    sec = np.array(fx["block"]["sectors"], np.int32)
    N, S2, d = (np.ascontiguousarray(sec[:, i]) for i in range(3))
    sites = np.array(fx["block"]["sites"], np.int32)
    ln = fx["block"]["length"]
    p = C.c_void_p()
    rc = lib.hh_plan_matvec(C.byref(mp), ln, len(N), L.ip(N), L.ip(S2), L.ip(d), len(sites), L.ip(sites), ln, len(N),
                            L.ip(N), L.ip(S2), L.ip(d), len(sites), L.ip(sites), fx["model"]["n_elec"], C.byref(p))
    assert rc == 0, lib.hh_last_error(None)
    try:
        out = {}
        for side in (0, 1):
            ns, nsub = L.i32(), L.i32()
            assert lib.hh_plan_enlarged(p, side, C.byref(ns), None, None, None, C.byref(nsub), None) == 0
            eN, eS, ed = (np.zeros(ns.value, np.int32) for _ in range(3))
            subs = np.zeros(5 * nsub.value, np.int32)
            lib.hh_plan_enlarged(p, side, C.byref(ns), L.ip(eN), L.ip(eS), L.ip(ed), C.byref(nsub), L.ip(subs))
            out[side] = (np.stack([eN, eS, ed], 1).tolist(), subs.reshape(-1, 5).tolist())
        dim, npairs = L.i64(), L.i32()
        lib.hh_plan_super(p, C.byref(dim), C.byref(npairs), None, None, None, None, None)
        cols = [np.zeros(npairs.value, np.int32) for _ in range(4)]
        off = np.zeros(npairs.value, np.int64)
        lib.hh_plan_super(p, C.byref(dim), C.byref(npairs), *[L.ip(c) for c in cols], L.ip(off))
        pairs = [[int(a), int(b), int(r), int(c), int(o)] for a, b, r, c, o in zip(*cols, off)]
        return out, pairs, dim.value
    finally:
        lib.hh_plan_destroy(p)


def test_product_basis_order_matches_hand_derived_fixture():
    fx = _fixture()
    out, pairs, dim = _product(fx)
    for side in (0, 1):
        sec, subs = out[side]
        assert sec == fx["enlarged_sectors"], side
        assert subs == fx["subsectors"], side
    assert pairs == fx["super_pairs"]
    assert dim == fx["super_dim"]


def test_oracle_basis_order_matches_hand_derived_fixture():
    fx = _fixture()
    bl = [(n, s) for n, s, _ in fx["block"]["sectors"]]
    dims = [d for _, _, d in fx["block"]["sectors"]]
    eb = dmrg.enlarge_basis(bl, dims, dmrg.site_sectors("hubbard"))
    got = [[n, s, d] for (n, s), d in zip(eb.sectors, eb.dims)]
    assert got == fx["enlarged_sectors"]
    subs = [[a, b, s, p, d] for a, lst in enumerate(eb.subs) for (b, s, p, d) in lst]
    assert subs == fx["subsectors"]
    sb = dmrg.super_basis(eb, eb, fx["model"]["n_elec"])
    assert sb.dim == fx["super_dim"]
    assert [list(p) for p in sb.pairs] == [q[:2] for q in fx["super_pairs"]]


def test_swapped_sort_keys_do_not_match_the_fixture():
    """A planted mistake: ordering N before 2S (or ascending) yields another
    sector order, which the fixture rejects."""
    fx = _fixture()
    exp = [tuple(x[:2]) for x in fx["enlarged_sectors"]]
    uniq = sorted(set(exp))
    assert sorted(uniq, key=lambda q: (-q[0], -q[1])) != exp      # N first
    assert sorted(uniq, key=lambda q: (q[1], q[0])) != exp        # ascending
    assert sorted(uniq, key=lambda q: (-q[1], -q[0])) == exp      # the reference's keys
