"""Pin of the oracle's superblock H (oracle/dmrg.py) against exact
diagonalisation: the full spectrum of the untruncated reduced superblock
operator equals the S = 0 spectrum of an independent Jordan-Wigner
Hamiltonian (oracle/ed.py)."""
import numpy as np
import pytest

from oracle import dmrg, ed
from oracle import lattice as lat


def _sb_matrix(model, ns, ne, N, keep=None):
    sg = dmrg.StepGeometry(model, ns, ne)
    if keep is not None:
        sg.bonds = {k: v for k, v in sg.bonds.items() if k in keep}
    se = dmrg.enlarge_in_geometry(dmrg.site_block(model, 0), model, "sys", sg)
    ee = dmrg.enlarge_in_geometry(dmrg.site_block(model, 0), model, "env", sg)
    H = dmrg.Superblock(se, ee, sg, model, N)
    D = H.sb.dim
    M = np.zeros((D, D), complex)
    for i in range(D):
        e = np.zeros(D, complex)
        e[i] = 1
        M[:, i] = H.matvec(e)
    return M, sg


def _ed(model, sg, N):
    g = sg.g
    pos = list(range(1, g.ns + 2)) + list(range(g.env_centre, g.totsite + 1))
    present = set(pos)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds if u in present and v in present]
    return ed.singlet_spectrum(model, pos, bonds, N)


CASES = [("hubbard", 0.0, 2), ("hubbard", 0.0, 4), ("hubbard", 0.125, 4), ("hubbard", 0.3, 3 + 1),
         ("tj", 0.0, 2), ("tj", 0.0, 4), ("tj", 0.2, 2)]


@pytest.mark.parametrize("kind,flux,N", CASES)
def test_superblock_spectrum_equals_ed(kind, flux, N):
    m = lat.Model(kind=kind, Lx=2, Ly=2, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    M, sg = _sb_matrix(m, 1, 1, N)
    assert np.abs(M - M.conj().T).max() < 1e-13
    ev = np.linalg.eigvalsh(M)
    ref = _ed(m, sg, N)
    assert len(ev) == len(ref)
    assert np.allclose(np.sort(ev), ref, atol=1e-10)


@pytest.mark.parametrize("kind", ["hubbard", "tj"])
@pytest.mark.parametrize("bond", [(1, 2), (3, 4), (1, 3), (2, 4)])
def test_each_bond_class_alone(kind, bond):
    # (1,2) sys block-site, (3,4) env site-block, (1,3) sys block-env site,
    # (2,4) sys site-env block: every term type of model_wave separately
    for N in (2, 4):
        m = lat.Model(kind=kind, Lx=2, Ly=2, U=8.0, J=0.4, t=1.0, flux=0.0, n_elec=N)
        M, sg = _sb_matrix(m, 1, 1, N, keep=[bond])
        assert np.allclose(np.sort(np.linalg.eigvalsh(M)), _ed(m, sg, N), atol=1e-10)


def test_larger_superblock_with_flux_cylinder():
    # 3x3 cylinder geometry with 2-site blocks: block-block bonds and the
    # periodic y wrap appear
    m = lat.Model(kind="tj", Lx=3, Ly=3, J=0.4, t=1.0, flux=1.0 / 9, n_elec=4)
    sg = dmrg.StepGeometry(m, 2, 2)
    sysb = dmrg.site_block(m, 0)
    envb = dmrg.site_block(m, 0)
    # grow both blocks once without truncation inside the small geometry
    sg1 = dmrg.StepGeometry(m, 1, 1)
    se = dmrg.enlarge_in_geometry(sysb, m, "sys", sg1)
    ee = dmrg.enlarge_in_geometry(envb, m, "env", sg1)
    sysb2 = dmrg.truncate_block(se, dmrg.identity_trun(se.basis), m, 2)
    # env block 2 must carry the bonds of the 3x3 geometry: rebuild in sg
    sg_e = dmrg.StepGeometry(m, 2, 2)
    ee2 = dmrg.enlarge_in_geometry(envb, m, "env", sg_e)
    envb2 = dmrg.truncate_block(ee2, dmrg.identity_trun(ee2.basis), m, 2)
    se3 = dmrg.enlarge_in_geometry(sysb2, m, "sys", sg)
    ee3 = dmrg.enlarge_in_geometry(envb2, m, "env", sg)
    H = dmrg.Superblock(se3, ee3, sg, m, 4)
    D = H.sb.dim
    M = np.zeros((D, D), complex)
    for i in range(D):
        e = np.zeros(D, complex)
        e[i] = 1
        M[:, i] = H.matvec(e)
    g = sg.g
    pos = list(range(1, g.ns + 2)) + list(range(g.env_centre, g.totsite + 1))
    present = set(pos)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds if u in present and v in present]
    ref = ed.singlet_spectrum(m, pos, bonds, 4, k=3)
    ev = np.sort(np.linalg.eigvalsh(M))
    assert np.allclose(ev[:3], ref[:3], atol=1e-9)
