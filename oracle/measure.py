"""Site densities <n_i> of a DMRG ground state -- TEST INFRASTRUCTURE ONLY.

Restates ``measure_operator_dia`` (code/measure.f90:3-61) with the operator
chain of ``Get_sys_oper_dia`` / ``Get_env_oper_dia``: the density of site i
enters as the new-site operator of the block of length i and is carried to
the measured block through every later truncation,

    O_i(l) = U_l^dag  embed_block(O_i(l-1))  U_l ,   O_i(i) = U_i^dag embed_site(n) U_i,

then read off the superblock ground state.  For a rank-0 block operator the
expectation reduces to Tr(rho_b O_b) / sqrt(2 S_b + 1) summed over the block
sectors b, with rho_b the block's reduced density matrix; the two centre
sites read sum_(b,s) n_s |psi_(b,s)|^2.

``cross_density`` gives the density-density correlations <n_i n_j> of the
pairs of ``Get_density_cor`` (code/correlation.f90:43-195) that straddle the
superblock centre (i on the system side, j on the environment side, centre
sites included): <psi| A_i (x) B_j |psi> = <psi, Y_i B_j^T> per superblock
pair block, Y_i = A_i psi, with the physical action O_b / sqrt(2 S_b + 1) of
a rank-0 reduced block operator on block sector b.

``spin_cross`` gives <S_i . S_j> for the same pairs across the centre: the
spin of every block site carried as a rank-1 operator (U_bra^dag O U_ket),
embedded in the superblock's enlarged bases and contracted as the exchange
bond, -sqrt(3) [S_i (x) S_j]^0.  ``hopping_cross`` gives sum_sigma
<c^dag_i c_j> the same way, with c^dag carried as a rank-1/2 fermionic
operator and the environment one read as its annihilator partner (the
hopping bond without its h.c.).  ``pair_cross`` gives the singlet pair
correlations <Delta^dag_a Delta_b> of bonds in the two blocks
(Get_mSCOP_Operator_cor0's block-block part), with each pair creator formed
as the Edmonds rank-0 product of the carried c^dag and the new site's c^dag.

``spin_row`` gives one row <S_r . S_j> over every j: block pairs through the
carried products S_r . S_j (the Edmonds rank-0 product of the carried spin
and the new site's spin), the block site with the system centre through the
same product in the superblock's enlarged basis, the environment side through
``spin_cross``.

``hopping_row`` gives one row sum_sigma <c^dag_r c_j> the same way, with
the fermionic products sqrt(2) [c^dag_a x c~_b]^0.

``density_row`` gives one row <n_r n_j> of the same correlation for a row
site r on the system side and every j, with the two-site products n_r n_j of
block sites carried through the truncations (``tracked_pairs``).

Pinned by exact diagonalisation (``ed.singlet_ground_density``,
``ed.singlet_ground_correlations``, ``ed.singlet_ground_spin_correlations``,
``ed.singlet_ground_hopping``, ``ed.singlet_ground_pairing``):
with no truncation the profile and the correlations equal the exact ones; the densities sum to N and, for each i,
sum_j <n_i n_j> over the other side equals <n_i N_other>; a full row plus
<n_r^2> sums to N <n_r>.
"""
import math

import numpy as np

from . import dmrg
from .su2 import phase


def _site_density_op(kind):
    return dmrg.site_ops(kind)["n"]


def _transform(op_enl, trun):
    """U^dag O U for a rank-0 enlarged operator; trun = [(enlarged sector, U)]."""
    out = dmrg.Op(0, 0, False)
    for i, (a, U) in enumerate(trun):
        m = op_enl.get((a, a))
        out.blk[(i, i)] = U.conj().T @ m @ U if m is not None else np.zeros((U.shape[1], U.shape[1]), complex)
    return out


def tracked_ops(model, blocks, truns, length, side):
    """{site k: O_k in the basis of blocks[length]} for k = 1 .. length."""
    st = dmrg.site_sectors(model.kind)
    n = _site_density_op(model.kind)
    ops = {1: n}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        new = {}
        for k, o in ops.items():
            new[k] = _transform(dmrg.embed_block_op(o, prev, eb, side), truns[l])
        new[l] = _transform(dmrg.embed_site_op(n, prev, eb, side), truns[l])
        ops = new
    return ops


def _block_rho(H, psi, side):
    """rho_b over the block sectors of `side` (sum over pairs and site states),
    and the centre-site density sum_(b,s) n_s |psi_(b,s)|^2."""
    sb = H.sb
    pin = dmrg.psi_blocks(sb, psi)
    en = H.s if side == "sys" else H.e
    eb = en.basis
    nsite = [st[0] for st in eb.st_sectors]   # N of each site multiplet
    rho = {}
    site_n = 0.0
    for i, (x, y) in enumerate(sb.pairs):
        a = x if side == "sys" else y
        for (b, s, pos, d) in eb.subs[a]:
            blk = pin[i][pos:pos + d, :] if side == "sys" else pin[i][:, pos:pos + d]
            # sys: rho_b = psi psi^dag; env: rho_b[c, c'] = sum_r psi[r, c] conj(psi[r, c'])
            r = blk @ blk.conj().T if side == "sys" else blk.T @ blk.conj()
            rho[b] = rho.get(b, 0.0) + r
            site_n += nsite[s] * float(np.real(np.vdot(blk, blk)))
    return rho, site_n


def density_profile(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H):
    """<n> at every lattice position (index position-1) for the ground state
    psi of the superblock H built from sys_blocks[sys_len], env_blocks[env_len]."""
    nsite = model.nsite
    out = np.zeros(nsite)
    for side, blocks, truns, length in (("sys", sys_blocks, sys_truns, sys_len),
                                        ("env", env_blocks, env_truns, env_len)):
        ops = tracked_ops(model, blocks, truns, length, side)
        rho, site_n = _block_rho(H, psi, side)
        secs = blocks[length].sectors
        for k, o in ops.items():
            v = 0.0
            for b, r in rho.items():
                m = o.blk.get((b, b))
                if m is not None:
                    v += float(np.real(np.trace(r @ m))) / math.sqrt(secs[b][1] + 1)
            pos = k if side == "sys" else nsite + 1 - k
            out[pos - 1] = v
        centre = length + 1 if side == "sys" else nsite - length
        out[centre - 1] = site_n
    return out


def _env_column_blocks(H, side):
    en = H.s if side == "sys" else H.e
    return en.basis, [st[0] for st in en.basis.st_sectors]


def cross_density(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H):
    """<n_i n_j> as an nsite x nsite array (index position - 1), symmetric,
    for i on the system side and j on the environment side; NaN elsewhere."""
    nsite = model.nsite
    out = np.full((nsite, nsite), np.nan)
    sb = H.sb
    pin = dmrg.psi_blocks(sb, psi)
    seb, s_n = _env_column_blocks(H, "sys")
    eeb, e_n = _env_column_blocks(H, "env")
    s_ops = tracked_ops(model, sys_blocks, sys_truns, sys_len, "sys")
    e_ops = tracked_ops(model, env_blocks, env_truns, env_len, "env")
    s_secs, e_secs = sys_blocks[sys_len].sectors, env_blocks[env_len].sectors
    s_list = [(k, s_ops[k]) for k in range(1, sys_len + 1)] + [(sys_len + 1, None)]
    e_list = [(k, e_ops[k]) for k in range(1, env_len + 1)] + [(env_len + 1, None)]
    for ki, A in s_list:
        # Y = (A (x) 1) psi, pair block by pair block
        Y = []
        for i, (x, _) in enumerate(sb.pairs):
            y = np.zeros_like(pin[i])
            for (b, s, pos, d) in seb.subs[x]:
                if A is None:
                    y[pos:pos + d, :] = s_n[s] * pin[i][pos:pos + d, :]
                else:
                    m = A.blk.get((b, b))
                    if m is not None:
                        y[pos:pos + d, :] = (m / math.sqrt(s_secs[b][1] + 1)) @ pin[i][pos:pos + d, :]
            Y.append(y)
        pi = ki
        for kj, B in e_list:
            v = 0.0
            for i, (_, yy) in enumerate(sb.pairs):
                for (b, s, pos, d) in eeb.subs[yy]:
                    ps, ys = pin[i][:, pos:pos + d], Y[i][:, pos:pos + d]
                    if B is None:
                        v += e_n[s] * float(np.real(np.vdot(ps, ys)))
                    else:
                        m = B.blk.get((b, b))
                        if m is not None:
                            v += float(np.real(np.vdot(ps, ys @ (m / math.sqrt(e_secs[b][1] + 1)).T)))
            pj = nsite + 1 - kj if B is not None else nsite - env_len
            out[pi - 1, pj - 1] = out[pj - 1, pi - 1] = v
    return out


def _product(A, B, eb):
    """Reduced matrices of the product of two rank-0 operators of the
    enlarged basis: (AB)_a = A_a B_a / sqrt(2 J_a + 1)."""
    out = {}
    for key, m in A.items():
        o = B.get(key)
        if o is not None:
            out[key] = m @ o / math.sqrt(eb.sectors[key[0]][1] + 1)
    return out


def tracked_pairs(model, blocks, truns, length, r, side="sys"):
    """{j: P_rj} with P_rj = n_r n_j (j != r, both block sites) in the basis of
    blocks[length] (Get_sys2b, code/correlation.f90:1-40): the product is
    formed in the enlarged basis of the step where its later site joins and is
    carried through every later truncation."""
    st = dmrg.site_sectors(model.kind)
    n = _site_density_op(model.kind)
    dens = {1: n}
    pairs = {}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        site = dmrg.embed_site_op(n, prev, eb, side)
        emb = lambda o: dmrg.embed_block_op(o, prev, eb, side)
        newp = {j: _transform(emb(P), truns[l]) for j, P in pairs.items()}
        if l < r:
            newd = {k: _transform(emb(o), truns[l]) for k, o in dens.items()}
            newd[l] = _transform(site, truns[l])
        elif l == r:
            for k, o in dens.items():
                newp[k] = _transform(_product(emb(o), site, eb), truns[l])
            newd = {r: _transform(site, truns[l])}
        else:
            e = emb(dens[r])
            newp[l] = _transform(_product(e, site, eb), truns[l])
            newd = {r: _transform(e, truns[l])}
        dens, pairs = newd, newp
    return pairs


def density_row(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H, r):
    """<n_r n_j> for every lattice position j (index j-1, NaN at j = r), for a
    row site r on the system side (1 .. sys_len + 1): one row of
    ``Get_density_cor`` (code/correlation.f90:43-195) over every j.

    Block-block pairs: Tr(rho_b P_b) / sqrt(2 S_b + 1) with the carried
    products P of ``tracked_pairs``.  Block site with the system centre site:
    sum over pair blocks and (b, s) of n_s <psi_sub, A_b psi_sub> /
    sqrt(2 S_b + 1).  Environment side: ``cross_density``."""
    nsite = model.nsite
    out = np.full(nsite, np.nan)
    cc = cross_density(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H)
    out[sys_len + 1:] = cc[r - 1, sys_len + 1:]
    ops = tracked_ops(model, sys_blocks, sys_truns, sys_len, "sys")
    secs = sys_blocks[sys_len].sectors
    pin = dmrg.psi_blocks(H.sb, psi)
    eb = H.s.basis
    ns = [st[0] for st in eb.st_sectors]

    def with_centre(A):
        v = 0.0
        for i, (x, _) in enumerate(H.sb.pairs):
            for (b, s, pos, d) in eb.subs[x]:
                m = A.blk.get((b, b))
                if m is None or ns[s] == 0:
                    continue
                blk = pin[i][pos:pos + d, :]
                v += ns[s] * float(np.real(np.vdot(blk, (m / math.sqrt(secs[b][1] + 1)) @ blk)))
        return v

    if r == sys_len + 1:
        for k in range(1, sys_len + 1):
            out[k - 1] = with_centre(ops[k])
        return out
    out[sys_len] = with_centre(ops[r])
    rho, _ = _block_rho(H, psi, "sys")
    for j, P in tracked_pairs(model, sys_blocks, sys_truns, sys_len, r).items():
        v = 0.0
        for b, rb in rho.items():
            m = P.blk.get((b, b))
            if m is not None:
                v += float(np.real(np.trace(rb @ m))) / math.sqrt(secs[b][1] + 1)
        out[j - 1] = v
    return out


def _transform_any(op_enl, trun, k2, dN=0, fermion=False):
    """U_bra^dag O U_ket for an enlarged operator of rank k2/2; trun =
    [(enlarged sector, U)] in new-block order."""
    pos = {a: i for i, (a, _) in enumerate(trun)}
    out = dmrg.Op(k2, dN, fermion)
    for (ap, a), m in op_enl.items():
        if ap in pos and a in pos:
            Ub, Uk = trun[pos[ap]][1], trun[pos[a]][1]
            out.blk[(pos[ap], pos[a])] = Ub.conj().T @ m @ Uk
    return out


def tracked_rank_ops(model, blocks, truns, length, side, name):
    """{site k: O_k in the basis of blocks[length]} for the site operator
    ``name`` ("S": rank 1; "cdag": rank 1/2, fermionic), carried through every
    truncation (Get_sys_oper_ndia / update_trun_ndia)."""
    st = dmrg.site_sectors(model.kind)
    O = dmrg.site_ops(model.kind)[name]
    ops = {1: O}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        tr = lambda e: _transform_any(e, truns[l], O.k2, O.dN, O.fermion)
        new = {k: tr(dmrg.embed_block_op(o, prev, eb, side)) for k, o in ops.items()}
        new[l] = tr(dmrg.embed_site_op(O, prev, eb, side))
        ops = new
    return ops


def tracked_spin_ops(model, blocks, truns, length, side):
    """{site k: S_k in the basis of blocks[length]}."""
    return tracked_rank_ops(model, blocks, truns, length, side, "S")


def _cross_apply(H, T, U, k2, amp, psi, fermion=False):
    """[T (x) U]^0 psi for enlarged-basis operators T (system) and U
    (environment), the cross-term contraction of the superblock
    (code/mat_vec.f90:1505-1928 without the h.c.); a fermionic U passes the
    system part, (-1)^{N_sys}."""
    sb = H.sb
    out = np.zeros_like(psi)
    pin = dmrg.psi_blocks(sb, psi)
    pout = dmrg.psi_blocks(sb, out)
    ssec = H.s.basis.sectors
    for (xp, x), tm in T.items():
        if x not in H.sidx or xp not in H.sidx:
            continue
        i, ip = H.sidx[x], H.sidx[xp]
        y, yp = sb.pairs[i][1], sb.pairs[ip][1]
        if (yp, y) not in U:
            continue
        J2, J2p = ssec[x][1], ssec[xp][1]
        c = amp / math.sqrt((k2 + 1) * (J2 + 1) * (J2p + 1))
        if fermion:
            c *= (-1.0) ** ssec[x][0]
        pout[ip] += c * (tm @ pin[i] @ U[(yp, y)].T)
    return out


def spin_cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H):
    """<S_i . S_j> = -sqrt(3) <[S_i (x) S_j]^0> for i on the system side and j
    on the environment side (nsite x nsite, symmetric, NaN elsewhere): the
    operators are embedded in the enlarged bases of the superblock and
    contracted as the superblock's exchange bond."""
    nsite = model.nsite
    out = np.full((nsite, nsite), np.nan)
    S = dmrg.site_ops(model.kind)["S"]
    s_ops = tracked_spin_ops(model, sys_blocks, sys_truns, sys_len, "sys")
    e_ops = tracked_spin_ops(model, env_blocks, env_truns, env_len, "env")
    sbl, ebl = sys_blocks[sys_len], env_blocks[env_len]
    seb, eeb = H.s.basis, H.e.basis
    T = {k: dmrg.embed_block_op(o, sbl, seb, "sys") for k, o in s_ops.items()}
    T[sys_len + 1] = dmrg.embed_site_op(S, sbl, seb, "sys")
    U = {k: dmrg.embed_block_op(o, ebl, eeb, "env") for k, o in e_ops.items()}
    U[env_len + 1] = dmrg.embed_site_op(S, ebl, eeb, "env")
    for ki, t in T.items():
        for kj, u in U.items():
            v = float(np.real(np.vdot(psi, _cross_apply(H, t, u, 2, -math.sqrt(3.0), psi))))
            pj = nsite + 1 - kj if kj <= env_len else nsite - env_len
            out[ki - 1, pj - 1] = out[pj - 1, ki - 1] = v
    return out


def hopping_cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H):
    """sum_sigma <c^dag_i,sigma c_j,sigma> = sqrt(2) <[c^dag_i (x) c~_j]^0> for i
    on the system side and j on the environment side (complex nsite x nsite;
    entry (j, i) the conjugate; NaN elsewhere): c^dag carried as a rank-1/2
    fermionic operator, the environment one read as its annihilator partner
    c~ (the superblock's hopping bond without its h.c.)."""
    nsite = model.nsite
    out = np.full((nsite, nsite), np.nan, dtype=complex)
    C = dmrg.site_ops(model.kind)["cdag"]
    s_ops = tracked_rank_ops(model, sys_blocks, sys_truns, sys_len, "sys", "cdag")
    e_ops = tracked_rank_ops(model, env_blocks, env_truns, env_len, "env", "cdag")
    sbl, ebl = sys_blocks[sys_len], env_blocks[env_len]
    seb, eeb = H.s.basis, H.e.basis
    T = {k: dmrg.embed_block_op(o, sbl, seb, "sys") for k, o in s_ops.items()}
    T[sys_len + 1] = dmrg.embed_site_op(C, sbl, seb, "sys")
    U = {k: dmrg.embed_block_op(o, ebl, eeb, "env") for k, o in e_ops.items()}
    U[env_len + 1] = dmrg.embed_site_op(C, ebl, eeb, "env")
    U = {k: dmrg._tilde_dict(u, eeb.sectors) for k, u in U.items()}
    for ki, t in T.items():
        for kj, u in U.items():
            v = complex(np.vdot(psi, _cross_apply(H, t, u, 1, math.sqrt(2.0), psi, fermion=True)))
            pj = nsite + 1 - kj if kj <= env_len else nsite - env_len
            out[ki - 1, pj - 1] = v
            out[pj - 1, ki - 1] = np.conj(v)
    return out


def _couple0(T, U, eb, k2):
    """[T x U]^0 of two rank-k2/2 operators on one enlarged basis (the
    operator product of block_mul_block_ndia): Edmonds (7.1.1) with k = 0,
    <a'||[T x U]^0||a> = (-1)^{2J} sum_{a''} (-1)^{k + J + J''} /
    sqrt((2k+1)(2J+1)) <a'||T||a''><a''||U||a>, J' = J."""
    out = {}
    for (ap, am), t in T.items():
        for (am2, a), u in U.items():
            if am2 != am or eb.sectors[ap][1] != eb.sectors[a][1]:
                continue
            J2, Jm2 = eb.sectors[a][1], eb.sectors[am][1]
            c = (-1.0) ** J2 * phase(k2 + J2 + Jm2) / math.sqrt((k2 + 1) * (J2 + 1))
            out[(ap, a)] = out.get((ap, a), 0.0) + c * (t @ u)
    return out


def tracked_pairs_cdag(model, blocks, truns, length, side, bonds):
    """{(s1, s2): P_12 in the basis of blocks[length]} for every bond (s1 <
    s2, block-site labels) with P_12 = [c^dag_s1 x c^dag_s2]^0, the singlet
    pair creator of Get_sys_two (code/correlation.f90:331-367): formed in the
    enlarged basis of the step where s2 joins, from the carried c^dag_s1 and
    the new site's c^dag, then carried through every later truncation as a
    rank-0, N + 2 operator."""
    st = dmrg.site_sectors(model.kind)
    C = dmrg.site_ops(model.kind)["cdag"]
    cd = {1: C}
    pairs = {}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        newp = {b: _transform_any(dmrg.embed_block_op(P, prev, eb, side), truns[l], 0, 2, False)
                for b, P in pairs.items()}
        site = dmrg.embed_site_op(C, prev, eb, side)
        for (s1, s2) in bonds:
            if s2 == l:
                P = _couple0(dmrg.embed_block_op(cd[s1], prev, eb, side), site, eb, 1)
                newp[(s1, s2)] = _transform_any(P, truns[l], 0, 2, False)
        cd = {k: _transform_any(dmrg.embed_block_op(o, prev, eb, side), truns[l], 1, 1, True) for k, o in cd.items()}
        cd[l] = _transform_any(site, truns[l], 1, 1, True)
        pairs = newp
    return pairs


def pair_cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H, sys_bonds,
               env_bonds):
    """<Delta^dag_b Delta_b'> for bonds b of the system block and b' of the
    environment block (block-site labels), Delta^dag_ij = (c^dag_i,up
    c^dag_j,down - c^dag_i,down c^dag_j,up) / sqrt(2): the block-block part of
    Get_mSCOP_Operator_cor0 (code/correlation.f90:198-327).  The environment
    annihilator is the adjoint of its carried creator; both are embedded in
    the superblock's enlarged bases and contracted as a rank-0 cross term.
    Returns a complex array (len(sys_bonds), len(env_bonds))."""
    Ps = tracked_pairs_cdag(model, sys_blocks, sys_truns, sys_len, "sys", sys_bonds)
    Pe = tracked_pairs_cdag(model, env_blocks, env_truns, env_len, "env", env_bonds)
    sbl, ebl = sys_blocks[sys_len], env_blocks[env_len]
    out = np.zeros((len(sys_bonds), len(env_bonds)), complex)
    for a, bs in enumerate(sys_bonds):
        T = dmrg.embed_block_op(Ps[bs], sbl, H.s.basis, "sys")
        for b, be in enumerate(env_bonds):
            Ue = dmrg.embed_block_op(Pe[be], ebl, H.e.basis, "env")
            U = {(y, yp): np.conj(m).T for (yp, y), m in Ue.items()}
            out[a, b] = complex(np.vdot(psi, _cross_apply(H, T, U, 0, 1.0, psi)))
    return out


def tracked_spin_products(model, blocks, truns, length, r, side="sys"):
    """{j: Q_rj} with Q_rj = S_r . S_j = -sqrt(3) [S_a x S_b]^0 (a < b the two
    sites) in the basis of blocks[length], formed as the Edmonds rank-0
    product of the carried spin of the earlier site and the new site's spin
    in the enlarged basis of the step where the later site joins
    (Get_sys2b), then carried as rank-0 operators.  Also returns the carried
    spins {k: S_k}.  r = 0 forms no products."""
    st = dmrg.site_sectors(model.kind)
    S = dmrg.site_ops(model.kind)["S"]
    spins = {1: S}
    prods = {}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        site = dmrg.embed_site_op(S, prev, eb, side)
        newp = {j: _transform(dmrg.embed_block_op(Q, prev, eb, side), truns[l]) for j, Q in prods.items()}
        created = []
        if l == r:
            created = [(j, spins[j]) for j in range(1, r)]
        elif 1 <= r < l:
            created = [(l, spins[r])]
        for j, Sj in created:
            Q = _couple0(dmrg.embed_block_op(Sj, prev, eb, side), site, eb, 2)
            newp[j] = _transform({k: -math.sqrt(3.0) * v for k, v in Q.items()}, truns[l])
        spins = {k: _transform_any(dmrg.embed_block_op(o, prev, eb, side), truns[l], 2) for k, o in spins.items()}
        spins[l] = _transform_any(site, truns[l], 2)
        prods = newp
    return prods, spins


def spin_row(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H, r):
    """<S_r . S_j> for every lattice position j (index j-1, NaN at j = r),
    row site r on the system side: one row of ``Get_density_cor`` with the
    spin.  Block pairs: Tr(rho_b Q_b) / sqrt(2 S_b + 1) with the carried
    products; block site with the system centre: the enlarged-basis product
    -sqrt(3) [S_r x S_centre]^0 read off psi; environment side:
    ``spin_cross``."""
    nsite = model.nsite
    out = np.full(nsite, np.nan)
    cc = spin_cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H)
    out[sys_len + 1:] = cc[r - 1, sys_len + 1:]
    prods, spins = tracked_spin_products(model, sys_blocks, sys_truns, sys_len, r if r <= sys_len else 0)
    sbl = sys_blocks[sys_len]
    eb = H.s.basis
    S = dmrg.site_ops(model.kind)["S"]
    site = dmrg.embed_site_op(S, sbl, eb, "sys")
    pin = dmrg.psi_blocks(H.sb, psi)

    def with_centre(Sj):
        Q = _couple0(dmrg.embed_block_op(Sj, sbl, eb, "sys"), site, eb, 2)
        v = 0.0
        for i, (x, _) in enumerate(H.sb.pairs):
            m = Q.get((x, x))
            if m is not None:
                v += -math.sqrt(3.0) * float(np.real(np.vdot(pin[i], m @ pin[i]))) / math.sqrt(eb.sectors[x][1] + 1)
        return v

    if r == sys_len + 1:
        for j in range(1, sys_len + 1):
            out[j - 1] = with_centre(spins[j])
        return out
    out[sys_len] = with_centre(spins[r])
    rho, _ = _block_rho(H, psi, "sys")
    secs = sbl.sectors
    for j, Q in prods.items():
        v = 0.0
        for b, rb in rho.items():
            m = Q.blk.get((b, b))
            if m is not None:
                v += float(np.real(np.trace(rb @ m))) / math.sqrt(secs[b][1] + 1)
        out[j - 1] = v
    return out


def hopping_row(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H, r):
    """sum_sigma <c^dag_r c_j> for every lattice position j (complex, index
    j-1, NaN at j = r), row site r on the system side.  Block pairs a < b:
    sum_sigma c^dag_a c_b = sqrt(2) [c^dag_a x c~_b]^0 formed as the Edmonds
    rank-0 product of the carried c^dag_a and the new site's annihilator
    partner when b joins, carried as a rank-0 operator (the row entry of
    j < r is the conjugate); block site with the system centre: the same
    product in the superblock's enlarged basis; environment side:
    ``hopping_cross``."""
    nsite = model.nsite
    out = np.full(nsite, np.nan, dtype=complex)
    cc = hopping_cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H)
    out[sys_len + 1:] = cc[r - 1, sys_len + 1:]
    st = dmrg.site_sectors(model.kind)
    C = dmrg.site_ops(model.kind)["cdag"]
    cd = {1: C}
    prods = {}
    for l in range(2, sys_len + 1):
        prev = sys_blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        site = dmrg.embed_site_op(C, prev, eb, "sys")
        site_t = dmrg._tilde_dict(site, eb.sectors)
        newp = {k: _transform(dmrg.embed_block_op(Q, prev, eb, "sys"), sys_truns[l]) for k, Q in prods.items()}
        created = []
        if r <= sys_len:
            if l == r:
                created = [((j, r), cd[j]) for j in range(1, r)]
            elif r < l:
                created = [((r, l), cd[r])]
        for key, Ca in created:
            Q = _couple0(dmrg.embed_block_op(Ca, prev, eb, "sys"), site_t, eb, 1)
            newp[key] = _transform({k: math.sqrt(2.0) * v for k, v in Q.items()}, sys_truns[l])
        cd = {k: _transform_any(dmrg.embed_block_op(o, prev, eb, "sys"), sys_truns[l], 1, 1, True)
              for k, o in cd.items()}
        cd[l] = _transform_any(site, sys_truns[l], 1, 1, True)
        prods = newp
    sbl = sys_blocks[sys_len]
    eb = H.s.basis
    site_t = dmrg._tilde_dict(dmrg.embed_site_op(C, sbl, eb, "sys"), eb.sectors)
    pin = dmrg.psi_blocks(H.sb, psi)

    def with_centre(Ca):
        Q = _couple0(dmrg.embed_block_op(Ca, sbl, eb, "sys"), site_t, eb, 1)
        v = 0.0
        for i, (x, _) in enumerate(H.sb.pairs):
            m = Q.get((x, x))
            if m is not None:
                v += math.sqrt(2.0) * complex(np.vdot(pin[i], m @ pin[i])) / math.sqrt(eb.sectors[x][1] + 1)
        return v

    if r == sys_len + 1:
        for j in range(1, sys_len + 1):
            out[j - 1] = np.conj(with_centre(cd[j]))
        return out
    out[sys_len] = with_centre(cd[r])
    rho, _ = _block_rho(H, psi, "sys")
    secs = sbl.sectors
    for (a, b), Q in prods.items():
        v = 0.0
        for blk, rb in rho.items():
            m = Q.blk.get((blk, blk))
            if m is not None:
                v += complex(np.trace(rb @ m)) / math.sqrt(secs[blk][1] + 1)
        if a == r:
            out[b - 1] = v
        else:
            out[a - 1] = np.conj(v)
    return out


def _scalar_times(P, U, eb):
    """(P U) for a rank-0 operator P (any N shift) and an operator U of any
    rank on one enlarged basis: P acts as P_red / sqrt(2 J + 1) on the
    multiplet, so (P U)_red[a', a] = sum_a'' P_red[a', a''] U_red[a'', a] /
    sqrt(2 J_a' + 1)."""
    out = {}
    for (ap, am), p in P.items():
        for (am2, a), u in U.items():
            if am2 != am:
                continue
            c = 1.0 / math.sqrt(eb.sectors[ap][1] + 1)
            out[(ap, a)] = out.get((ap, a), 0.0) + c * (p @ u)
    return out


def pair_same_block(model, sys_blocks, sys_truns, sys_len, psi, H, quads):
    """<Delta^dag_12 Delta_34> for bonds (s1, s2) and (s3, s4) of the system
    block with s2 < s3 < s4 (Get_sys_two1, code/correlation.f90:370-428,
    measured as measure_sys_block): the pair creator P_12 is formed when s2
    joins; at s3 it is multiplied by the new site's annihilator partner
    (a rank-1/2, N + 1 operator W = P_12 c~_3), at s4 W is coupled with the
    next site's annihilator partner to the scalar Z = [W x c~_4]^0 =
    -P_12 Delta_34 (the sign of the annihilator tensor convention), and Z is
    carried to the measured block and read off its reduced density matrix.
    Returns a complex array over `quads`."""
    st = dmrg.site_sectors(model.kind)
    C = dmrg.site_ops(model.kind)["cdag"]
    cd = {1: C}
    P, W, Z = {}, {}, {}
    for l in range(2, sys_len + 1):
        prev = sys_blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        tr = sys_truns[l]
        site = dmrg.embed_site_op(C, prev, eb, "sys")
        site_t = dmrg._tilde_dict(site, eb.sectors)
        emb = lambda o: dmrg.embed_block_op(o, prev, eb, "sys")
        nP = {k: _transform_any(emb(o), tr, 0, 2, False) for k, o in P.items()}
        nW = {k: _transform_any(emb(o), tr, 1, 1, True) for k, o in W.items()}
        nZ = {k: _transform(emb(o), tr) for k, o in Z.items()}
        for q, ((s1, s2), (s3, s4)) in enumerate(quads):
            if s2 == l:
                nP[q] = _transform_any(_couple0(emb(cd[s1]), site, eb, 1), tr, 0, 2, False)
            if s3 == l:
                nW[q] = _transform_any(_scalar_times(emb(P[q]), site_t, eb), tr, 1, 1, True)
            if s4 == l:
                nZ[q] = _transform({k: -v for k, v in _couple0(emb(W[q]), site_t, eb, 1).items()}, tr)
        cd = {k: _transform_any(emb(o), tr, 1, 1, True) for k, o in cd.items()}
        cd[l] = _transform_any(site, tr, 1, 1, True)
        P, W, Z = nP, nW, nZ
    rho, _ = _block_rho(H, psi, "sys")
    secs = sys_blocks[sys_len].sectors
    out = np.zeros(len(quads), complex)
    for q in range(len(quads)):
        v = 0.0
        for b, rb in rho.items():
            m = Z[q].blk.get((b, b))
            if m is not None:
                v += complex(np.trace(rb @ m)) / math.sqrt(secs[b][1] + 1)
        out[q] = v
    return out


def _onsite_op(kind, name):
    """Rank-0 site operator whose expectation is the on-site entry of
    Get_oper_cor_ndia1 (st_oper x st_oper1^dag of one site): n^2 for the
    density, S.S = S(S+1) for the spin, sum_sigma c^dag c = n for the
    one-body correlation; reduced elements sqrt(2S+1) f(s)."""
    op = dmrg.Op(0, 0, False)
    for i, (N, S2) in enumerate(dmrg.site_sectors(kind)):
        f = {"n": N * N, "S": S2 * (S2 + 2) / 4.0, "cdag": N}[name]
        if f:
            op.add((i, i), [[math.sqrt(S2 + 1) * f]])
    return op


def _carry_site_op(model, blocks, truns, length, side, site_op):
    """{k: f_k in the basis of blocks[length]} for a rank-0 site operator."""
    st = dmrg.site_sectors(model.kind)
    ops = {1: site_op}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        new = {k: _transform(dmrg.embed_block_op(o, prev, eb, side), truns[l]) for k, o in ops.items()}
        new[l] = _transform(dmrg.embed_site_op(site_op, prev, eb, side), truns[l])
        ops = new
    return ops


def _expect_block(rho, op, secs):
    """<O> = sum_b Tr(rho_b O_b) / sqrt(2 S_b + 1) for a rank-0 block operator."""
    v = 0j
    for b, r in rho.items():
        m = op.blk.get((b, b))
        if m is not None:
            v += complex(np.trace(r @ m)) / math.sqrt(secs[b][1] + 1)
    return v


def _pair_product(name, e, site, eb):
    """Two-site scalar of block operator e and new-site operator site, both in
    one enlarged basis: n_a n_b, S_a . S_b = -sqrt(3) [S_a x S_b]^0 or
    sum_sigma c^dag_a c_b = sqrt(2) [c^dag_a x c~_b]^0."""
    if name == "n":
        return _product(e, site, eb)
    if name == "S":
        return {k: -math.sqrt(3.0) * v for k, v in _couple0(e, site, eb, 2).items()}
    return {k: math.sqrt(2.0) * v for k, v in _couple0(e, dmrg._tilde_dict(site, eb.sectors), eb, 1).items()}


def tracked_all_products(model, blocks, truns, length, side, name):
    """({(a, b): Q_ab}, {k: O_k}) in the basis of blocks[length] for every pair
    a < b of block sites of one side (Get_sys_oper_cor_ndia1 /
    Get_env_oper_cor_ndia1, code/correlation.f90:851-974, :1097-1216): Q_ab
    is formed in the enlarged basis of the step where b joins, from the
    carried O_a and the new site's operator (``_pair_product``), then carried
    as a rank-0 operator; O_k is the carried single-site operator."""
    st = dmrg.site_sectors(model.kind)
    O = dmrg.site_ops(model.kind)[name]
    ones = {1: O}
    prods = {}
    for l in range(2, length + 1):
        prev = blocks[l - 1]
        eb = dmrg.enlarge_basis(prev.sectors, prev.dims, st)
        site = dmrg.embed_site_op(O, prev, eb, side)
        emb = {k: dmrg.embed_block_op(o, prev, eb, side) for k, o in ones.items()}
        newp = {key: _transform(dmrg.embed_block_op(Q, prev, eb, side), truns[l]) for key, Q in prods.items()}
        for a, e in emb.items():
            newp[(a, l)] = _transform(_pair_product(name, e, site, eb), truns[l])
        if name == "n":
            ones = {k: _transform(e, truns[l]) for k, e in emb.items()}
            ones[l] = _transform(site, truns[l])
        else:
            ones = {k: _transform_any(e, truns[l], O.k2, O.dN, O.fermion) for k, e in emb.items()}
            ones[l] = _transform_any(site, truns[l], O.k2, O.dN, O.fermion)
        prods = newp
    return prods, ones


def corr_all(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H, name):
    """Get_oper_cor_ndia1 (code/correlation.f90:2891-3107) for name "n"
    (<n_i n_j>), "S" (<S_i . S_j>) or "cdag" (sum_sigma <c^dag_i c_j>): every
    pair of lattice positions and the on-site values, (nsite, nsite), index
    position - 1; real for "n" and "S", complex Hermitian for "cdag".

    Pairs across the centre: ``cross_density`` / ``spin_cross`` /
    ``hopping_cross``.  Pairs of block sites of one side: the products of
    ``tracked_all_products`` read off that side's block density matrix.  A
    block site with its side's centre site: the same two-site scalar formed
    in the superblock's enlarged basis and applied to psi (rows of the system
    part, columns of the environment part).  On site: the carried
    ``_onsite_op`` of every block site, and sum f(s) |psi_(b,s)|^2 for the
    centre sites."""
    nsite = model.nsite
    out = np.full((nsite, nsite), np.nan, dtype=complex)
    cross = {"n": cross_density, "S": spin_cross, "cdag": hopping_cross}[name]
    cc = cross(model, sys_blocks, env_blocks, sys_truns, env_truns, sys_len, env_len, psi, H)
    fin = ~np.isnan(cc)
    out[fin] = cc[fin]
    pin = dmrg.psi_blocks(H.sb, psi)
    O = dmrg.site_ops(model.kind)[name]
    fop = _onsite_op(model.kind, name)
    fval = [fop.blk[(i, i)][0, 0] / math.sqrt(S2 + 1) if (i, i) in fop.blk else 0.0
            for i, (_, S2) in enumerate(dmrg.site_sectors(model.kind))]
    for side, blocks, truns, length in (("sys", sys_blocks, sys_truns, sys_len),
                                        ("env", env_blocks, env_truns, env_len)):
        pos = (lambda k: k) if side == "sys" else (lambda k, n=nsite: n + 1 - k)
        centre = sys_len + 1 if side == "sys" else nsite - env_len
        blk = blocks[length]
        secs = blk.sectors
        rho, _ = _block_rho(H, psi, side)
        prods, ones = tracked_all_products(model, blocks, truns, length, side, name)
        for (a, b), Q in prods.items():
            v = _expect_block(rho, Q, secs)
            out[pos(a) - 1, pos(b) - 1] = v
            out[pos(b) - 1, pos(a) - 1] = np.conj(v)
        for k, f in _carry_site_op(model, blocks, truns, length, side, fop).items():
            out[pos(k) - 1, pos(k) - 1] = _expect_block(rho, f, secs).real
        eb = H.s.basis if side == "sys" else H.e.basis
        site = dmrg.embed_site_op(O, blk, eb, side)
        diag = 0.0
        for i, (x, y) in enumerate(H.sb.pairs):
            z = x if side == "sys" else y
            for (b, s, p0, d) in eb.subs[z]:
                part = pin[i][p0:p0 + d, :] if side == "sys" else pin[i][:, p0:p0 + d]
                diag += fval[s] * float(np.real(np.vdot(part, part)))
        out[centre - 1, centre - 1] = diag
        for a, o in ones.items():
            Q = _pair_product(name, dmrg.embed_block_op(o, blk, eb, side), site, eb)
            v = 0j
            for i, (x, y) in enumerate(H.sb.pairs):
                z = x if side == "sys" else y
                m = Q.get((z, z))
                if m is None:
                    continue
                w = pin[i]
                v += complex(np.vdot(w, m @ w if side == "sys" else w @ m.T)) / math.sqrt(eb.sectors[z][1] + 1)
            out[pos(a) - 1, centre - 1] = v
            out[centre - 1, pos(a) - 1] = np.conj(v)
    return out if name == "cdag" else out.real
