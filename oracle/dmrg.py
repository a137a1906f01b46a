"""SU(2) two-site DMRG, numpy restatement of the reference inner loop.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The product never imports
this module.

Conventions (identical to the reference unless noted):
  * quantum numbers are (N, 2S); a block sector list is ordered 2S descending
    then N descending when built by ``enlarge_basis`` (``Get_basis`` +
    ``QuickSort_Integer``, code/update.f90:11-112, :240-273); enlarged
    subsectors are block-major / site-minor with offsets ``spos``;
  * operators are reduced tensors in Edmonds' convention, i.e. the identity
    has reduced element sqrt(2S+1) and scalar operators are stored as
    sqrt(2S+1) x matrix (code/warm_up.f90:818-988, code/update.f90:460,
    code/mat_vec.f90:193); site elements c^dag = -sqrt2 (0->1/2),
    sqrt2 (1/2->double), n = (0, sqrt2, 2), double = (0,0,1), S = sqrt(3/2);
  * superblock [sys_bl x sys_st] x [env_st x env_bl], fermion order equal to
    the lattice path, psi stored per (sys sector, env sector) pair as a
    column-major sys_dim x env_dim matrix (``Sub_Wave``, code/pubdata.f90:126);
  * Lanczos uses real-part inner products and classical Gram-Schmidt against
    all previous vectors (code/lanczos.f90:4-176, code/space_info.f90:1018-1214);
  * truncation keeps between kept_min and kept_max multiplets by discarded
    weight (code/get_den.f90:184-428).
Deliberate difference: env-side reduced matrices hold the physical operator
(the reference stores their complex conjugate and applies psi.B^dag); here
env operators act as psi.B^T, the env density matrix is psi^T conj(psi) and
the env truncation is applied with U^T.  Gauge-invariant outputs (energies,
kept counts, truncation errors) are unaffected.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import lattice as lat
from .su2 import w6j, phase

SQ2 = math.sqrt(2.0)


# ----------------------------------------------------------------------------
# operators and blocks
# ----------------------------------------------------------------------------
class Op:
    """Reduced tensor operator: blk[(bra, ket)] = matrix (dim_bra x dim_ket)."""

    def __init__(self, k2, dN, fermion):
        self.k2 = k2
        self.dN = dN
        self.fermion = fermion
        self.blk = {}

    def add(self, key, mat):
        if key in self.blk:
            self.blk[key] = self.blk[key] + mat
        else:
            self.blk[key] = np.array(mat, dtype=np.complex128)


def site_sectors(kind):
    return [(0, 0), (1, 1), (2, 0)] if kind == "hubbard" else [(0, 0), (1, 1)]


def site_ops(kind):
    """Single-site reduced operators (``Get_site_operator_su2``,
    code/warm_up.f90:818-991)."""
    secs = site_sectors(kind)
    idx = {q: i for i, q in enumerate(secs)}
    cd = Op(1, 1, True)
    cd.add((idx[(1, 1)], idx[(0, 0)]), [[-SQ2]])
    if kind == "hubbard":
        cd.add((idx[(2, 0)], idx[(1, 1)]), [[SQ2]])
    n = Op(0, 0, False)
    n.add((idx[(1, 1)], idx[(1, 1)]), [[SQ2]])
    if kind == "hubbard":
        n.add((idx[(2, 0)], idx[(2, 0)]), [[2.0]])
    dbl = Op(0, 0, False)
    if kind == "hubbard":
        dbl.add((idx[(2, 0)], idx[(2, 0)]), [[1.0]])
    S = Op(2, 0, False)
    S.add((idx[(1, 1)], idx[(1, 1)]), [[math.sqrt(1.5)]])
    return {"cdag": cd, "n": n, "double": dbl, "S": S}


@dataclass
class Block:
    length: int
    sectors: list
    dims: list
    H: Op
    ops: dict = field(default_factory=dict)   # block site index -> {name: Op}

    def index(self):
        return {q: i for i, q in enumerate(self.sectors)}


def site_block(model, site_label_U):
    """Length-1 block: H = U*double (code/warm_up.f90:1301-1325)."""
    secs = site_sectors(model.kind)
    so = site_ops(model.kind)
    H = Op(0, 0, False)
    for i, _ in enumerate(secs):
        H.add((i, i), [[0.0]])
    if model.kind == "hubbard" and model.U != 0.0:
        for (a, b), m in so["double"].blk.items():
            H.blk[(a, b)] = H.blk[(a, b)] + model.U * m
    ops = {1: {nm: so[nm] for nm in _op_names(model)}}
    return Block(1, list(secs), [1] * len(secs), H, ops)


def _op_names(model):
    return ["cdag"] if model.kind == "hubbard" else ["cdag", "S", "n"]


# ----------------------------------------------------------------------------
# enlarged basis (Get_basis) and super basis (Get_super_basis)
# ----------------------------------------------------------------------------
@dataclass
class EnlBasis:
    sectors: list
    dims: list
    subs: list          # per sector: list of (b, s, spos, sdim)
    bl_sectors: list
    st_sectors: list

    def index(self):
        return {q: i for i, q in enumerate(self.sectors)}


def enlarge_basis(bl_sectors, bl_dims, st_sectors):
    """``Get_basis`` (code/update.f90:11-112)."""
    pre = set()
    for (Nb, Sb) in bl_sectors:
        for (Ns, Ss) in st_sectors:
            for S in range(abs(Sb - Ss), Sb + Ss + 1, 2):
                pre.add((Nb + Ns, S))
    secs = sorted(pre, key=lambda q: (-q[1], -q[0]))
    subs, dims = [], []
    for (N, S) in secs:
        lst, pos = [], 0
        for b, (Nb, Sb) in enumerate(bl_sectors):
            for s, (Ns, Ss) in enumerate(st_sectors):
                if Nb + Ns == N and abs(Sb - Ss) <= S <= Sb + Ss:
                    lst.append((b, s, pos, bl_dims[b]))
                    pos += bl_dims[b]
        subs.append(lst)
        dims.append(pos)
    return EnlBasis(secs, dims, subs, list(bl_sectors), list(st_sectors))


@dataclass
class SuperBasis:
    pairs: list          # (sys sector idx, env sector idx)
    offsets: list
    shapes: list
    dim: int


def super_basis(sys_bs, env_bs, N, S2=0):
    """``Get_super_basis`` (code/update.f90:118-171); sys-major order."""
    pairs, offs, shapes, off = [], [], [], 0
    for x, (Nx, Sx) in enumerate(sys_bs.sectors):
        for y, (Ny_, Sy) in enumerate(env_bs.sectors):
            if Ny_ == N - Nx and abs(Sx - Sy) <= S2 <= Sx + Sy:
                pairs.append((x, y))
                offs.append(off)
                shapes.append((sys_bs.dims[x], env_bs.dims[y]))
                off += sys_bs.dims[x] * env_bs.dims[y]
    return SuperBasis(pairs, offs, shapes, off)


def psi_blocks(sb, vec):
    """Views of the flat psi as column-major sector matrices."""
    return [vec[o:o + r * c].reshape((r, c), order="F") for o, (r, c) in zip(sb.offsets, sb.shapes)]


# ----------------------------------------------------------------------------
# SU(2) coefficients
# ----------------------------------------------------------------------------
def e1(b2p, b2, s2, J2p, J2, k2):
    """<(b' s)J'||T(1)||(b s)J> / <b'||T||b>   (Edmonds 7.1.7)."""
    return phase(b2p + s2 + J2 + k2) * math.sqrt((J2 + 1) * (J2p + 1)) * w6j(b2p, J2p, s2, J2, b2, k2)


def e2(s2p, s2, b2, J2p, J2, k2):
    """<(b s')J'||U(2)||(b s)J> / <s'||U||s>   (Edmonds 7.1.8)."""
    return phase(b2 + s2 + J2p + k2) * math.sqrt((J2 + 1) * (J2p + 1)) * w6j(s2p, J2p, b2, J2, s2, k2)


def x12(b2p, b2, s2p, s2, J2, k2):
    """Matrix element factor of [T(1) x U(2)]^0 on (b s)J (reduced / sqrt(2J+1))."""
    return phase(b2 + k2 + s2p + J2) / math.sqrt(k2 + 1) * w6j(b2p, b2, k2, s2, s2p, J2)


def adjoint_tilde(op, sectors):
    """Reduced matrices of the annihilator-type partner a~ of a creator-type
    rank-1/2 tensor (a~_q = (-1)^(1/2+q) a_-q):
    <j||a~||j'> = -(-1)^(1/2+j'-j) conj(<j'||c^dag||j>)  (block_transfer_trans,
    code/space_info.f90:32-63)."""
    out = Op(op.k2, -op.dN, op.fermion)
    for (bra, ket), m in op.blk.items():
        j2p = sectors[bra][1]
        j2 = sectors[ket][1]
        sgn = -phase(op.k2 + j2p - j2)
        out.add((ket, bra), sgn * np.conj(m.T))
    return out


def adjoint_vector(op, sectors):
    """Partner of a Hermitian rank-k tensor (S): S~ = -S for k=1, so the
    scalar coupling [S x S]^0 uses S itself; returned unchanged."""
    return op


# ----------------------------------------------------------------------------
# enlarged blocks (block + one site)
# ----------------------------------------------------------------------------
@dataclass
class Enlarged:
    side: str               # "sys" | "env"
    basis: EnlBasis
    H: dict                 # sector -> reduced matrix (scalar)
    ops: dict               # site key -> {name: {(a', a): reduced matrix}}, keys are
                            # ("bl", k) for block site k and ("st", 0) for the new site
    k2: dict = field(default_factory=dict)


def _nsign(N):
    return -1.0 if N % 2 else 1.0


def _sub_index(eb):
    """(block sector, site sector) -> [(enlarged sector, spos, sdim), ...];
    one block x site pair appears in every allowed total spin."""
    idx = {}
    for a, lst in enumerate(eb.subs):
        for (b, s, pos, d) in lst:
            idx.setdefault((b, s), []).append((a, pos, d))
    return idx


def embed_block_op(op, bl, eb, side):
    """Block operator -> enlarged basis (``update_block_ndia_sys/env``,
    code/update.f90:548-963).  Env blocks sit after their site in the fermion
    order, so a fermionic block operator picks up (-1)^{N_site}."""
    out = {}
    idx = _sub_index(eb)
    for (bp, b), m in op.blk.items():
        for s, (Ns, Ss) in enumerate(eb.st_sectors):
            for (a, pos, d) in idx.get((b, s), ()):
                for (ap, posp, dp) in idx.get((bp, s), ()):
                    J2, J2p = eb.sectors[a][1], eb.sectors[ap][1]
                    c = e1(bl.sectors[bp][1], bl.sectors[b][1], Ss, J2p, J2, op.k2)
                    if c == 0.0:
                        continue
                    if side == "env" and op.fermion:
                        c *= _nsign(Ns)
                    key = (ap, a)
                    if key not in out:
                        out[key] = np.zeros((eb.dims[ap], eb.dims[a]), np.complex128)
                    out[key][posp:posp + dp, pos:pos + d] += c * m
    return out


def embed_site_op(op, bl, eb, side):
    """New-site operator -> enlarged basis (``update_site_ndia_sys/env``,
    code/update.f90:969-1381).  On the sys side the site follows the block in
    the fermion order: (-1)^{N_block}."""
    out = {}
    idx = _sub_index(eb)
    for (sp, s), m in op.blk.items():
        for b, (Nb, Sb) in enumerate(bl.sectors):
            for (a, pos, d) in idx.get((b, s), ()):
                for (ap, posp, dp) in idx.get((b, sp), ()):
                    J2, J2p = eb.sectors[a][1], eb.sectors[ap][1]
                    c = e2(eb.st_sectors[sp][1], eb.st_sectors[s][1], Sb, J2p, J2, op.k2)
                    if c == 0.0:
                        continue
                    if side == "sys" and op.fermion:
                        c *= _nsign(Nb)
                    key = (ap, a)
                    if key not in out:
                        out[key] = np.zeros((eb.dims[ap], eb.dims[a]), np.complex128)
                    out[key][posp:posp + dp, pos:pos + d] += c * m[0, 0] * np.eye(d)
    return out


def pair_scalar(Tb, Us, bl, eb, side, k2, fermion):
    """Reduced matrices of [T(block) x U(site)]^0 in the enlarged basis, with
    the operator product written in fermion order (sys: T_bl U_st; env:
    T'_st U'_bl, passed here as Tb = block factor, Us = site factor).

    sys: sign (-1)^{N_block,in} (U acts first and passes the block);
    env: product T_st U_bl, U acts first and passes the site,
         sign (-1)^{N_site,in}; [T(2) x U(1)]^0 = (-1)^{2k} [U(1) x T(2)]^0.
    Returns {a: reduced matrix} (scalar, diagonal in the enlarged sector).
    """
    out = {}
    idx = _sub_index(eb)
    for (bp, b), mb in Tb.blk.items():
        for (sp, s), ms in Us.blk.items():
          for (a, pos, d) in idx.get((b, s), ()):
            for (ap, posp, dp) in idx.get((bp, sp), ()):
              if ap != a:
                continue
              J2 = eb.sectors[a][1]
              c = x12(bl.sectors[bp][1], bl.sectors[b][1], eb.st_sectors[sp][1],
                      eb.st_sectors[s][1], J2, k2)
              if c == 0.0:
                  continue
              if fermion:
                  c *= _nsign(bl.sectors[b][0]) if side == "sys" else _nsign(eb.st_sectors[s][0])
              if side == "env" and (k2 % 2 == 1):
                  c = -c
              c *= math.sqrt(J2 + 1)       # reduced = sqrt(2J+1) * matrix
              if a not in out:
                  out[a] = np.zeros((eb.dims[a], eb.dims[a]), np.complex128)
              out[a][posp:posp + dp, pos:pos + d] += c * ms[0, 0] * mb
    return out


def enlarge(bl, model, side, bonds):
    """Block + one site, all operators on the enlarged basis.

    bl: Block; side: "sys"|"env"; bonds: list of (k, hop, J, V) couplings
    between block site k (which has stored operators) and the new site, with
    hop the amplitude of c^dag_first c_second in fermion order (sys: block
    site first; env: new site first).  Builds H (``Get_Hamiltonian_Sys/Env``,
    code/get_ham.f90:4-659 before its truncation) and the local operators of
    every block boundary site and of the new site.
    """
    st = site_sectors(model.kind)
    so = site_ops(model.kind)
    eb = enlarge_basis(bl.sectors, bl.dims, st)
    H = {}

    def acc(dst, src):
        for k, m in src.items():
            dst[k] = dst[k] + m if k in dst else m.copy()

    # embedded block H (update_block_dia) and on-site U (update_site_dia)
    hb = embed_block_op(bl.H, bl, eb, side)
    acc(H, {a: m for (ap, a), m in hb.items()})
    if model.kind == "hubbard" and model.U != 0.0:
        du = embed_site_op(so["double"], bl, eb, side)
        acc(H, {a: model.U * m for (ap, a), m in du.items()})
    for a in range(len(eb.sectors)):
        if a not in H:
            H[a] = np.zeros((eb.dims[a], eb.dims[a]), np.complex128)

    # bonds between block boundary sites and the new site
    for (k, amp_hop, jspin, vden) in bonds:
        ops_k = bl.ops[k]
        if amp_hop != 0.0:
            if side == "sys":
                # amp c^dag_k c_s  = amp sqrt2 [c^dag_k(1) x a~_s(2)]^0
                Tb = ops_k["cdag"]
                Us = adjoint_tilde(so["cdag"], st)
            else:
                # env order (site, block): amp c^dag_s c_k = amp sqrt2 [c^dag_s x a~_k]^0
                Tb = adjoint_tilde(ops_k["cdag"], bl.sectors)
                Us = so["cdag"]
            F = pair_scalar(Tb, Us, bl, eb, side, 1, True)
            for a, m in F.items():
                f = amp_hop * SQ2 * m
                H[a] = H[a] + f + f.conj().T
        if jspin != 0.0:
            F = pair_scalar(ops_k["S"], so["S"], bl, eb, side, 2, False)
            for a, m in F.items():
                H[a] = H[a] + (-math.sqrt(3.0) * jspin) * m
        if vden != 0.0:
            F = pair_scalar(ops_k["n"], so["n"], bl, eb, side, 0, False)
            for a, m in F.items():
                H[a] = H[a] + vden * m

    # local operators on the enlarged basis
    ops = {}
    for k, d in bl.ops.items():
        ops[("bl", k)] = {nm: embed_block_op(o, bl, eb, side) for nm, o in d.items()}
    ops[("st", 0)] = {nm: embed_site_op(so[nm], bl, eb, side) for nm in _op_names(model)}
    k2 = {"cdag": 1, "S": 2, "n": 0}
    return Enlarged(side, eb, H, ops, k2)


# ----------------------------------------------------------------------------
# geometry-aware helpers
# ----------------------------------------------------------------------------
class StepGeometry:
    """Positions and bonds of one superblock step (``dmrg_setup``)."""

    def __init__(self, model, ns, ne):
        self.model = model
        self.g = lat.geometry(model, ns, ne)
        self.bonds = {}
        for u, v, kind, amp in lat.lattice_bonds(model, self.g.ncol):
            self.bonds[(u, v)] = (kind, amp)

    def sys_pos(self, k):
        return k

    def env_pos(self, k):
        return self.g.totsite + 1 - k

    def coupling(self, pu, pv):
        """(hop amp of c^dag_pu c_pv, J, V) for positions pu, pv (or None)."""
        key = (min(pu, pv), max(pu, pv))
        if key not in self.bonds:
            return None
        kind, amp = self.bonds[key]
        if pu > pv:
            amp = amp.conjugate()
        m = self.model
        if m.kind == "hubbard":
            return (amp, 0.0, 0.0)
        return (amp, m.J, -0.25 * m.J)


def enlarge_in_geometry(bl, model, side, sg):
    """``enlarge`` with the block-to-new-site bonds of the step geometry sg."""
    pos = sg.sys_pos if side == "sys" else sg.env_pos
    pn = pos(bl.length + 1)
    bonds = []
    for k in bl.ops:
        c = sg.coupling(pos(k), pn) if side == "sys" else sg.coupling(pn, pos(k))
        if c is not None:
            bonds.append((k,) + c)
    return enlarge(bl, model, side, bonds)


# ----------------------------------------------------------------------------
# truncation of an enlarged block into a new block
# ----------------------------------------------------------------------------
def truncate_block(en, trun, model, new_length):
    """Renormalise H and the kept boundary operators with the truncation
    (``update_trun_dia/ndia``, code/update.f90:1387-1605).

    trun: list of (enlarged sector index, U) in the new block's sector order.
    """
    eb = en.basis
    new_secs = [eb.sectors[a] for a, _ in trun]
    new_dims = [U.shape[1] for _, U in trun]
    pos = {a: i for i, (a, _) in enumerate(trun)}
    Hn = Op(0, 0, False)
    for i, (a, U) in enumerate(trun):
        Hn.blk[(i, i)] = U.conj().T @ en.H[a] @ U
    keep = set(lat.boundary_sites(model, new_length))
    ops = {}
    for key, d in en.ops.items():
        k = key[1] if key[0] == "bl" else new_length
        if k not in keep:
            continue
        ops[k] = {}
        for nm, blk in d.items():
            o = Op(en.k2[nm], 1 if nm == "cdag" else 0, nm == "cdag")
            for (ap, a), m in blk.items():
                if ap in pos and a in pos:
                    o.blk[(pos[ap], pos[a])] = trun[pos[ap]][1].conj().T @ m @ trun[pos[a]][1]
            ops[k][nm] = o
    return Block(new_length, new_secs, new_dims, Hn, ops)


def identity_trun(eb):
    """Warm-up truncation while the enlarged dim <= kept_max
    (code/warm_up.f90:1421-1432): identity in Get_basis order."""
    return [(a, np.eye(d, dtype=np.complex128)) for a, d in enumerate(eb.dims)]


# ----------------------------------------------------------------------------
# superblock H psi
# ----------------------------------------------------------------------------
class Superblock:
    """Superblock operator built from two enlarged blocks (``hmerge`` +
    ``model_wave``, code/mat_vec.f90:3476-3852)."""

    def __init__(self, sys_en, env_en, sg, model, N):
        self.s, self.e = sys_en, env_en
        self.model = model
        self.sb = super_basis(sys_en.basis, env_en.basis, N)
        ssec = sys_en.basis.sectors
        esec = env_en.basis.sectors
        # env-side partners: a~ (hop) / S / n of every env site, merged per sys site
        pos_s = {("bl", k): sg.sys_pos(k) for (t, k) in sys_en.ops if t == "bl"}
        pos_s[("st", 0)] = sg.g.sys_centre
        pos_e = {("bl", k): sg.env_pos(k) for (t, k) in env_en.ops if t == "bl"}
        pos_e[("st", 0)] = sg.g.env_centre
        self.cross = []      # (name, k2, T (sys enl op dict), U (env enl op dict), amp, fermion)
        for ks, ps in pos_s.items():
            merged = {"cdag": {}, "S": {}, "n": {}}
            used = {"cdag": False, "S": False, "n": False}
            for ke, pe in pos_e.items():
                c = sg.coupling(ps, pe)
                if c is None:
                    continue
                hop, Jc, Vc = c
                if hop != 0.0:
                    at = _tilde_dict(env_en.ops[ke]["cdag"], esec)
                    _acc_scaled(merged["cdag"], at, hop)
                    used["cdag"] = True
                if Jc != 0.0:
                    _acc_scaled(merged["S"], env_en.ops[ke]["S"], Jc)
                    used["S"] = True
                if Vc != 0.0:
                    _acc_scaled(merged["n"], env_en.ops[ke]["n"], Vc)
                    used["n"] = True
            if used["cdag"]:
                self.cross.append(("hop", 1, sys_en.ops[ks]["cdag"], merged["cdag"], SQ2, True))
            if used["S"]:
                self.cross.append(("spin", 2, sys_en.ops[ks]["S"], merged["S"], -math.sqrt(3.0), False))
            if used["n"]:
                self.cross.append(("den", 0, sys_en.ops[ks]["n"], merged["n"], 1.0, False))
        self.sidx = {x: i for i, (x, y) in enumerate(self.sb.pairs)}

    def matvec(self, v):
        sb = self.sb
        out = np.zeros_like(v)
        pin = psi_blocks(sb, v)
        pout = psi_blocks(sb, out)
        ssec = self.s.basis.sectors
        esec = self.e.basis.sectors
        for i, (x, y) in enumerate(sb.pairs):
            J2 = ssec[x][1]
            pout[i] += (self.s.H[x] / math.sqrt(J2 + 1)) @ pin[i]
            pout[i] += pin[i] @ (self.e.H[y] / math.sqrt(J2 + 1)).T
        for (name, k2, T, U, amp, ferm) in self.cross:
            for (xp, x), tm in T.items():
                if x not in self.sidx or xp not in self.sidx:
                    continue
                i = self.sidx[x]
                ip = self.sidx[xp]
                y = sb.pairs[i][1]
                yp = sb.pairs[ip][1]
                if (yp, y) not in U:
                    continue
                J2, J2p = ssec[x][1], ssec[xp][1]
                c = amp / math.sqrt((k2 + 1) * (J2 + 1) * (J2p + 1))
                if ferm:
                    c *= _nsign(ssec[x][0])
                um = U[(yp, y)]
                pout[ip] += c * (tm @ pin[i] @ um.T)
                if name == "hop":
                    # h.c.: adjoint of the forward map
                    pout[i] += np.conj(c) * (tm.conj().T @ pin[ip] @ um.conj())
        return out


def _tilde_dict(blkdict, sectors):
    out = {}
    for (ap, a), m in blkdict.items():
        sgn = -phase(1 + sectors[ap][1] - sectors[a][1])
        out[(a, ap)] = sgn * np.conj(m.T)
    return out


def _acc_scaled(dst, src, c):
    for k, m in src.items():
        dst[k] = dst[k] + c * m if k in dst else c * m


# ----------------------------------------------------------------------------
# Lanczos (lanczos_memo_new)
# ----------------------------------------------------------------------------
def _rdot(a, b):
    """wave_product: real part of <a,b> (code/space_info.f90:1018-1034)."""
    return float(np.real(np.vdot(a, b)))


def tridiag_eigh(alpha, beta):
    """DSTEV on the Lanczos tridiagonal (ascending eigenpairs)."""
    n = len(alpha)
    T = np.diag(alpha)
    if n > 1:
        T += np.diag(beta[:n - 1], 1) + np.diag(beta[:n - 1], -1)
    return np.linalg.eigh(T)


def _deflate(w, lower):
    """w minus its complex projection on the (orthonormal) lower levels: the
    real-coefficient projection on u and on i u for every u."""
    for u in lower:
        w = w - _rdot(u, w) * u - _rdot(1j * u, w) * (1j * u)
    return w


def lanczos(matvec, v0, lanzero, max_iter=100, lower=()):
    """``lanczos_memo_new`` (code/lanczos.f90:4-176) for one level.

    lower: unit eigenvectors of the levels below (levels > 1): the start
    vector and every new Lanczos vector are projected off them
    (Gram_Schmidt_Lanczos, code/space_info.f90:1222-1279), by complex
    projection, and beta is the norm after the projection (Lanczos on the
    projected operator).  Returns (E, psi, iterations)."""
    v0 = _deflate(v0, lower)
    V = [v0 / math.sqrt(_rdot(v0, v0))]
    w = matvec(V[0])
    alpha = [_rdot(V[0], w)]
    beta = []
    ritz = alpha[0]
    last = 1
    while last < max_iter:
        if last < 2:
            w = w - alpha[last - 1] * V[last - 1]
        else:
            w = w - alpha[last - 1] * V[last - 1] - beta[last - 2] * V[last - 2]
        w = _deflate(w, lower)
        b = math.sqrt(_rdot(w, w))
        beta.append(b)
        if abs(b) <= lanzero:
            break
        w = w / b
        # classical Gram-Schmidt with real coefficients (Gram_Schmidt_process)
        proj = np.zeros_like(w)
        for u in V:
            proj += (_rdot(u, w) / _rdot(u, u)) * u
        w = w - proj
        w = w / math.sqrt(_rdot(w, w))
        V.append(w)
        last += 1
        w = matvec(V[-1])
        alpha.append(_rdot(V[-1], w))
        ev, _ = tridiag_eigh(np.array(alpha), np.array(beta))
        if abs(ev[0] - ritz) <= lanzero or last == max_iter:
            ritz = ev[0]
            break
        ritz = ev[0]
    ev, z = tridiag_eigh(np.array(alpha), np.array(beta + [0.0]))
    psi = np.zeros_like(V[0])
    for i in range(last):
        psi += z[i, 0] * V[i]
    psi /= math.sqrt(_rdot(psi, psi))
    return float(ev[0]), psi, last


def lanczos_levels(matvec, starts, lanzero, max_iter=100):
    """``lanczos_memo_new`` with levels = len(starts) (code/lanczos.f90:29-44):
    the levels one by one, level k from starts[k] projected off levels < k.

    The reference projects with real coefficients only (wave_product,
    Gram_Schmidt_Lanczos), which leaves i u of every lower level u; in the
    real inner-product space that is an eigenvector of energy E_u, so its
    level 2 returns E_0 again.  Here the projection is complex (u and i u).
    Returns (energies, vectors, iterations)."""
    es, vs, its = [], [], []
    for v0 in starts:
        e, psi, it = lanczos(matvec, v0, lanzero, max_iter, lower=vs)
        es.append(e)
        vs.append(psi)
        its.append(it)
    return es, vs, its


# ----------------------------------------------------------------------------
# density matrix and truncation (get_den.f90)
# ----------------------------------------------------------------------------
def density_matrix(sb, psi, side):
    """``Get_density_matrix_sys/env`` (code/get_den.f90:4-99): one block per
    psi sector, in super-basis order; returns list of (enlarged sector, rho)."""
    out = []
    for (x, y), m in zip(sb.pairs, psi_blocks(sb, psi)):
        if side == "sys":
            out.append((x, m @ m.conj().T))
        else:
            out.append((y, m.T @ m.conj()))
    return out


def density_matrix_levels(sb, psis, side):
    """``Get_density_matrix_sys_new`` / ``_env_new`` (code/get_den.f90:103-177):
    with one level the density matrix of psis[0]; with two or more,
    (0.1 rho(psis[0]) + 1.9 rho(psis[1])) / 2 (the other levels do not
    enter).  0.1 and 1.9 are default-real (single precision) literals in the
    reference (get_den.f90:133, :171), so the weights are float32(0.1) / 2 and
    float32(1.9) / 2 and the trace is 1 - 1.1e-8."""
    r0 = density_matrix(sb, psis[0], side)
    if len(psis) == 1:
        return r0
    r1 = density_matrix(sb, psis[1], side)
    w0, w1 = float(np.float32(0.1)), float(np.float32(1.9))
    return [(a, (w0 * m0 + w1 * m1) / 2.0) for (a, m0), (_, m1) in zip(r0, r1)]


@dataclass
class TruncResult:
    trun: list           # (enlarged sector, U) in new block order
    num_state: int
    err: float
    entropy: float
    kept_per_sector: list


def truncation(rho, kept_min, kept_max, trun_err):
    """``Get_truncation_operator`` (code/get_den.f90:184-428).

    Sort key: eigenvalue descending, ties broken by the position in the
    concatenated per-sector ascending spectra (deterministic stand-in for the
    unstable QuickSort_Double)."""
    evals, evecs, gidx = [], [], []
    off = 0
    for (a, r) in rho:
        w, v = np.linalg.eigh(r)
        evals.append(w)
        evecs.append(v)
        gidx.append(np.arange(off, off + len(w)))
        off += len(w)
    allv = np.concatenate(evals) if evals else np.zeros(0)
    alli = np.concatenate(gidx) if gidx else np.zeros(0, int)
    order = np.lexsort((alli, -allv))
    sv = allv[order]
    si = alli[order]
    tot = 0.0
    for x in sv:
        tot += x
    sv = sv / tot
    dim = len(sv)
    if dim <= kept_min:
        num, err = dim, 0.0
    else:
        err = 0.0
        for i in range(kept_min, dim):
            err += sv[i]
        num = kept_min
        while err > trun_err:
            num += 1
            err -= sv[num - 1]
            if err <= trun_err or num >= kept_max or num >= dim:
                break
    ent = 0.0
    for x in sv:
        if x >= 1.0e-15:
            ent -= x * math.log(x)
    kept_set = set(si[:num].tolist())
    trun, per = [], []
    for (a, r), w, v, gi in zip(rho, evals, evecs, gidx):
        c = sum(1 for g in gi if g in kept_set)
        per.append(c)
        if c > 0:
            trun.append((a, v[:, len(w) - c:]))
    return TruncResult(trun, num, float(err), ent, per)


# ----------------------------------------------------------------------------
# wave transformation between steps (wave_trans_LR / RL)
# ----------------------------------------------------------------------------
def wave_trans_LR(old_sb, old_sys_en, old_env_en, psi, sys_trun, env_trun_of_old_envblock,
                  new_sys_en, new_env_en, new_sb):
    """Predict psi for the step after moving the centre one site right
    (``wave_trans_LR``, code/warm_up.f90:1855-2012).

    old env enlarged = (E, e) with E = old env block (truncation of
    new_env_en with env_trun); new sys enlarged = (a, e), a = truncated old
    sys enlarged.  psi' = sign * U_s^dag psi U_e^T with the recoupling sign
    (-1)^{3 j_E + 3 j_e + J}."""
    out = np.zeros(new_sb.dim, np.complex128)
    pin = psi_blocks(old_sb, psi)
    pout = psi_blocks(new_sb, out)
    strun = {a: (i, U) for i, (a, U) in enumerate(sys_trun)}
    etrun = env_trun_of_old_envblock        # list (new_env enlarged sector, U) in E order
    new_pairs = {p: i for i, p in enumerate(new_sb.pairs)}
    nsidx = new_sys_en.basis.index()
    neidx = new_env_en.basis.index()
    oeb = old_env_en.basis
    # new sys enlarged subsector lookup: (a, site s) -> (sector, pos)
    nsub = {}
    for A, lst in enumerate(new_sys_en.basis.subs):
        for (b, s, p, d) in lst:
            nsub[(b, s, new_sys_en.basis.sectors[A][1])] = (A, p, d)
    for i, (x, y) in enumerate(old_sb.pairs):
        if x not in strun:
            continue
        ia, Us = strun[x]
        J2 = old_sys_en.basis.sectors[x][1]
        left = Us.conj().T @ pin[i]                   # kept_a x env_dim
        for (E, e, p, d) in oeb.subs[y]:
            # E: index into the old env block sectors; U_e maps new env enlarged
            encl, Ue = etrun[E]
            E2 = oeb.bl_sectors[E][1]
            e2 = oeb.st_sectors[e][1]
            if (ia, e, E2) not in nsub:
                continue
            A, pa, da = nsub[(ia, e, E2)]
            key = (A, encl)
            if key not in new_pairs:
                continue
            sgn = phase(3 * E2 + 3 * e2 + J2)
            pout[new_pairs[key]][pa:pa + da, :] += sgn * (left[:, p:p + d] @ Ue.T)
    n = math.sqrt(_rdot(out, out))
    return out / n if n > 0 else out


def wave_trans_RL(old_sb, old_sys_en, old_env_en, psi, sys_trun_of_old_sysblock, env_trun,
                  new_sys_en, new_env_en, new_sb):
    """Mirror of ``wave_trans_LR`` (``wave_trans_RL``, code/warm_up.f90:2016-2170).

    old sys enlarged = (A, s), A = old sys block = truncation of new_sys_en;
    new env enlarged = (f', s) with f' = truncated old env enlarged.
    psi' = sign * U_s psi conj(U_e) with sign (-1)^{j_s + J - j_A}."""
    out = np.zeros(new_sb.dim, np.complex128)
    pin = psi_blocks(old_sb, psi)
    pout = psi_blocks(new_sb, out)
    etrun = {a: (i, U) for i, (a, U) in enumerate(env_trun)}
    strun = sys_trun_of_old_sysblock
    new_pairs = {p: i for i, p in enumerate(new_sb.pairs)}
    osb = old_sys_en.basis
    nesub = {}
    for B, lst in enumerate(new_env_en.basis.subs):
        for (b, s, p, d) in lst:
            nesub[(b, s, new_env_en.basis.sectors[B][1])] = (B, p, d)
    for i, (x, y) in enumerate(old_sb.pairs):
        if y not in etrun:
            continue
        jf, Ue = etrun[y]
        J2 = old_env_en.basis.sectors[y][1]
        right = pin[i] @ Ue.conj()                   # sys_dim x kept_f
        for (A, s, p, d) in osb.subs[x]:
            encl, Us = strun[A]
            A2 = osb.bl_sectors[A][1]
            s2 = osb.st_sectors[s][1]
            if (jf, s, A2) not in nesub:
                continue
            B, pb, db = nesub[(jf, s, A2)]
            key = (encl, B)
            if key not in new_pairs:
                continue
            sgn = phase(s2 + J2 - A2)
            pout[new_pairs[key]][:, pb:pb + db] += sgn * (Us @ right[p:p + d, :])
    n = math.sqrt(_rdot(out, out))
    return out / n if n > 0 else out


# ----------------------------------------------------------------------------
# deterministic start vectors (counter-based, identical in the product)
# ----------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def start_vector(dim, seed):
    """Uniform [0,1) real and imaginary parts (as ``wave_initialize``,
    code/space_info.f90:1079-1105) from splitmix64(seed*2^32 + 2g (+1)),
    normalised."""
    g = np.arange(dim, dtype=np.uint64)
    base = np.uint64((seed & 0xFFFFFFFF) << 32)
    with np.errstate(over="ignore"):
        re = splitmix64(base + np.uint64(2) * g)
        im = splitmix64(base + np.uint64(2) * g + np.uint64(1))
    scale = 1.0 / float(1 << 53)
    v = (re >> np.uint64(11)).astype(np.float64) * scale + 1j * (im >> np.uint64(11)).astype(np.float64) * scale
    return v / math.sqrt(float(np.real(np.vdot(v, v))))
