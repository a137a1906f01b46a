"""Exact diagonalisation in the Fock basis (Jordan-Wigner), oracle pin.

TEST INFRASTRUCTURE ONLY.  Independent of every SU(2) routine: it builds the
Hamiltonian of the same lattice / couplings (oracle/lattice.py) on bit
strings, with modes ordered (site 1 up, site 1 down, site 2 up, ...) along the
DMRG path, restricted to fixed N and S_z = 0, and projects onto total S = 0
with the Casimir S^2.
"""
from itertools import combinations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _popcount(a):
    a = a.astype(np.int64)
    c = np.zeros_like(a)
    while np.any(a):
        c += a & 1
        a >>= 1
    return c


class FockSector:
    def __init__(self, nsite, n_elec, kind):
        self.nsite = nsite
        self.kind = kind
        states = []
        half = n_elec // 2
        for ups in combinations(range(nsite), half):
            su = sum(1 << (2 * i) for i in ups)
            for dns in combinations(range(nsite), n_elec - half):
                if kind == "tj" and set(ups) & set(dns):
                    continue
                states.append(su + sum(1 << (2 * i + 1) for i in dns))
        states.sort()
        self.states = np.array(states, dtype=np.int64)
        self.index = {int(s): i for i, s in enumerate(states)}

    def apply(self, ops, vec_states=None):
        """Apply a product of fermion operators (list of (mode, dagger)),
        rightmost first, to every basis state.  Returns (src_idx, dst_idx, sign)
        for the states where the result stays in the sector."""
        s = self.states.copy()
        sign = np.ones(len(s))
        ok = np.ones(len(s), dtype=bool)
        for mode, dag in reversed(ops):
            bit = (s >> mode) & 1
            if dag:
                ok &= bit == 0
            else:
                ok &= bit == 1
            below = s & ((1 << mode) - 1)
            sign *= np.where(_popcount(below) % 2 == 1, -1.0, 1.0)
            s = s ^ (1 << mode)
        src = np.nonzero(ok)[0]
        dst, keep = [], []
        for i in src:
            j = self.index.get(int(s[i]))
            if j is not None:
                dst.append(j)
                keep.append(i)
        keep = np.array(keep, dtype=np.int64)
        return keep, np.array(dst, dtype=np.int64), sign[keep]


def hamiltonian(model, positions, bonds, n_elec):
    """Sparse H on the ordered site list ``positions`` (lattice positions in
    fermion order) with ``bonds`` = list of (pu, pv, hop, J, V), hop = amp of
    c^dag_pu c_pv (h.c. added)."""
    order = {p: i for i, p in enumerate(positions)}
    fs = FockSector(len(positions), n_elec, model.kind)
    dim = len(fs.states)
    rows, cols, vals = [], [], []

    def add(ops, c):
        src, dst, sg = fs.apply(ops)
        rows.extend(dst.tolist())
        cols.extend(src.tolist())
        vals.extend((c * sg).tolist())

    for (pu, pv, hop, J, V) in bonds:
        i, j = order[pu], order[pv]
        if hop != 0:
            for sgm in (0, 1):
                add([(2 * i + sgm, True), (2 * j + sgm, False)], hop)
                add([(2 * j + sgm, True), (2 * i + sgm, False)], np.conj(hop))
        if J != 0:
            # S_i.S_j = Sz Sz + (S+_i S-_j + S-_i S+_j)/2
            for a in (0, 1):
                for b in (0, 1):
                    sa = 0.5 if a == 0 else -0.5
                    sb = 0.5 if b == 0 else -0.5
                    add([(2 * i + a, True), (2 * i + a, False), (2 * j + b, True), (2 * j + b, False)], J * sa * sb)
            add([(2 * i, True), (2 * i + 1, False), (2 * j + 1, True), (2 * j, False)], 0.5 * J)
            add([(2 * i + 1, True), (2 * i, False), (2 * j, True), (2 * j + 1, False)], 0.5 * J)
        if V != 0:
            for a in (0, 1):
                for b in (0, 1):
                    add([(2 * i + a, True), (2 * i + a, False), (2 * j + b, True), (2 * j + b, False)], V)
    if model.kind == "hubbard" and model.U != 0:
        for i in range(len(positions)):
            add([(2 * i, True), (2 * i, False), (2 * i + 1, True), (2 * i + 1, False)], model.U)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim), dtype=np.complex128)
    return H, fs


def casimir(fs):
    """Total S^2 on the S_z = 0 sector."""
    n = fs.nsite
    rows, cols, vals = [], [], []
    dim = len(fs.states)

    def add(ops, c):
        src, dst, sg = fs.apply(ops)
        rows.extend(dst.tolist())
        cols.extend(src.tolist())
        vals.extend((c * sg).tolist())

    for i in range(n):
        for j in range(n):
            for a in (0, 1):
                for b in (0, 1):
                    sa = 0.5 if a == 0 else -0.5
                    sb = 0.5 if b == 0 else -0.5
                    add([(2 * i + a, True), (2 * i + a, False), (2 * j + b, True), (2 * j + b, False)], sa * sb)
            add([(2 * i, True), (2 * i + 1, False), (2 * j + 1, True), (2 * j, False)], 0.5)
            add([(2 * i + 1, True), (2 * i, False), (2 * j, True), (2 * j + 1, False)], 0.5)
    return sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim), dtype=np.complex128)


def singlet_spectrum(model, positions, bonds, n_elec, k=None):
    """Eigenvalues of H restricted to total S = 0 (all of them when the
    sector is small, else the lowest k via H + 50 S^2)."""
    H, fs = hamiltonian(model, positions, bonds, n_elec)
    S2 = casimir(fs)
    dim = H.shape[0]
    if k is None and dim <= 3000:
        w, v = np.linalg.eigh(H.toarray())
        s2 = np.real(np.einsum("ij,ij->j", v.conj(), S2.toarray() @ v))
        # degenerate multiplets mix S values; diagonalise S2 within each
        # degenerate H eigenspace
        out = []
        i = 0
        while i < dim:
            j = i
            while j + 1 < dim and abs(w[j + 1] - w[i]) < 1e-8:
                j += 1
            blk = v[:, i:j + 1]
            ss = np.linalg.eigvalsh(blk.conj().T @ (S2 @ blk))
            out.extend([w[i]] * int(np.sum(np.abs(ss) < 1e-6)))
            i = j + 1
        return np.array(out)
    kk = k or 4
    w = spla.eigsh(H + 50.0 * S2, k=kk, which="SA", return_eigenvectors=False)
    return np.sort(np.real(w))


def _singlet_ground(model, positions, bonds, n_elec):
    H, fs = hamiltonian(model, positions, bonds, n_elec)
    S2 = casimir(fs)
    dim = H.shape[0]
    if dim <= 3000:
        w, v = np.linalg.eigh((H + 50.0 * S2).toarray())
    else:
        w, v = spla.eigsh(H + 50.0 * S2, k=1, which="SA")
    return float(np.real(w[0])), v[:, 0], fs


def _occupations(fs, positions):
    """occ[pos - 1] = occupation of that position in every Fock state."""
    occ = np.zeros((max(positions), len(fs.states)))
    for j, pos in enumerate(positions):
        occ[pos - 1] = ((fs.states >> (2 * j)) & 1) + ((fs.states >> (2 * j + 1)) & 1)
    return occ


def singlet_ground_density(model, positions, bonds, n_elec):
    """(E0, <n> per lattice position) of the lowest total-S = 0 state;
    the density array is indexed by position - 1."""
    e, psi, fs = _singlet_ground(model, positions, bonds, n_elec)
    return e, _occupations(fs, positions) @ (np.abs(psi) ** 2)


def singlet_ground_correlations(model, positions, bonds, n_elec):
    """(E0, <n_i>, <n_i n_j>) of the lowest total-S = 0 state, indexed by
    lattice position - 1.  The number operators are diagonal in the Fock
    basis, so <n_i n_j> = sum_f |psi_f|^2 occ_i(f) occ_j(f)."""
    e, psi, fs = _singlet_ground(model, positions, bonds, n_elec)
    occ = _occupations(fs, positions)
    p = np.abs(psi) ** 2
    return e, occ @ p, (occ * p) @ occ.T


def singlet_ground_spin_correlations(model, positions, bonds, n_elec):
    """(E0, <S_i . S_j>) of the lowest total-S = 0 state, indexed by lattice
    position - 1 (i != j; the diagonal is 3/4 <n_i - 2 n_i,up n_i,down>, left
    as computed by the same formula)."""
    e, psi, fs = _singlet_ground(model, positions, bonds, n_elec)
    n = max(positions)
    out = np.zeros((n, n))
    order = {p: i for i, p in enumerate(positions)}
    dim = len(fs.states)
    for pa in positions:
        for pb in positions:
            i, j = order[pa], order[pb]
            rows, cols, vals = [], [], []

            def add(ops, c):
                src, dst, sg = fs.apply(ops)
                rows.extend(dst.tolist())
                cols.extend(src.tolist())
                vals.extend((c * sg).tolist())

            for a in (0, 1):
                for b in (0, 1):
                    sa = 0.5 if a == 0 else -0.5
                    sb = 0.5 if b == 0 else -0.5
                    add([(2 * i + a, True), (2 * i + a, False), (2 * j + b, True), (2 * j + b, False)], sa * sb)
            add([(2 * i, True), (2 * i + 1, False), (2 * j + 1, True), (2 * j, False)], 0.5)
            add([(2 * i + 1, True), (2 * i, False), (2 * j, True), (2 * j + 1, False)], 0.5)
            O = sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim), dtype=np.complex128)
            out[pa - 1, pb - 1] = float(np.real(np.vdot(psi, O @ psi)))
    return e, out


def singlet_ground_hopping(model, positions, bonds, n_elec):
    """(E0, G) with G[i-1, j-1] = sum_sigma <c^dag_i,sigma c_j,sigma> in the
    lowest total-S = 0 state (lattice positions; the diagonal is <n_i>)."""
    e, psi, fs = _singlet_ground(model, positions, bonds, n_elec)
    n = max(positions)
    out = np.zeros((n, n), complex)
    order = {p: i for i, p in enumerate(positions)}
    for pa in positions:
        for pb in positions:
            i, j = order[pa], order[pb]
            v = 0.0
            for sgm in (0, 1):
                src, dst, sg = fs.apply([(2 * i + sgm, True), (2 * j + sgm, False)])
                v += np.sum(np.conj(psi[dst]) * sg * psi[src])
            out[pa - 1, pb - 1] = v
    return e, out


def singlet_ground_pairing(model, positions, bonds, n_elec, left, right):
    """(E0, <Delta^dag_b Delta_b'>) for bonds b in ``left`` and b' in
    ``right`` (pairs of lattice positions), Delta^dag_ij = (c^dag_i,up
    c^dag_j,down - c^dag_i,down c^dag_j,up) / sqrt(2), in the lowest total-S
    = 0 state."""
    e, psi, fs = _singlet_ground(model, positions, bonds, n_elec)
    order = {p: i for i, p in enumerate(positions)}
    out = np.zeros((len(left), len(right)), complex)
    for a, (pi, pj) in enumerate(left):
        for b, (pk, pl) in enumerate(right):
            i, j, k, l = order[pi], order[pj], order[pk], order[pl]
            v = 0.0
            # (1/2) sum of the four singlet products
            for (s1, s2, s3, s4, c) in ((0, 1, 1, 0, 1.0), (0, 1, 0, 1, -1.0), (1, 0, 1, 0, -1.0), (1, 0, 0, 1, 1.0)):
                ops = [(2 * i + s1, True), (2 * j + s2, True), (2 * l + s3, False), (2 * k + s4, False)]
                src, dst, sg = fs.apply(ops)
                v += 0.5 * c * np.sum(np.conj(psi[dst]) * sg * psi[src])
            out[a, b] = v
    return e, out
