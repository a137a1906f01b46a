"""Pins for the oracle's sweep machinery: with no truncation every finite
sweep step equals the exact ground state (growth, renormalisation, wave
transformation); Lanczos against numpy eigh; the kept-count rule on
synthetic spectra."""
import math

import numpy as np
import pytest

from oracle import dmrg, driver, ed
from oracle import lattice as lat


def _ed_gs(m):
    sg = dmrg.StepGeometry(m, m.nsite // 2 - 1, m.nsite - m.nsite // 2 - 1)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    return ed.singlet_spectrum(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)[0]


@pytest.mark.parametrize("kind,Lx,Ly,flux,N", [("hubbard", 4, 2, 0.15, 6), ("tj", 4, 2, 0.2, 6),
                                               ("hubbard", 3, 2, 0.0, 4)])
def test_untruncated_sweeps_equal_ed(kind, Lx, Ly, flux, N):
    m = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    e0 = _ed_gs(m)
    recs = driver.run(m, driver.Schedule(kept_max=2000, n_sweeps=1, tol=1e-12))
    sweep = [r for r in recs if r.phase == "sweep1"]
    assert sweep
    for r in sweep:
        assert abs(r.energy - e0) < 1e-10 * abs(e0), (r.sys_len, r.energy, e0)
        # the transformed guess is already the ground state
        assert r.iters <= 3


def test_wave_transformation_is_isometry_and_preserves_energy():
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=4.0, t=1.0, flux=0.1, n_elec=6)
    run = driver.Run(m, driver.Schedule(kept_max=2000, n_sweeps=0, tol=1e-12))
    run.run()
    # redo one LR step by hand from the stored state
    n = m.nsite
    label = 2
    sysb = run.sys_blocks[label]
    envb = run.env_blocks[n - label - 2]
    E, psi, it, se, ee, H = run._solve_with_guess(sysb, envb, m.n_elec, 1e-13, None, None, "LR")
    rho = dmrg.density_matrix(H.sb, psi, "sys")
    tr = dmrg.truncation(rho, 10 ** 6, 10 ** 6, 0.0)
    newsys = dmrg.truncate_block(se, tr.trun, m, label + 1)
    sg2 = dmrg.StepGeometry(m, label + 1, n - label - 3)
    se2 = dmrg.enlarge_in_geometry(newsys, m, "sys", sg2)
    ee2 = dmrg.enlarge_in_geometry(run.env_blocks[n - label - 3], m, "env", sg2)
    H2 = dmrg.Superblock(se2, ee2, sg2, m, m.n_elec)
    psi2 = dmrg.wave_trans_LR(H.sb, se, ee, psi, tr.trun, run.env_trun[n - label - 2], se2, ee2, H2.sb)
    assert abs(np.vdot(psi2, psi2).real - 1.0) < 1e-12
    e2 = np.vdot(psi2, H2.matvec(psi2)).real
    assert abs(e2 - E) < 1e-9


def test_lanczos_matches_eigh():
    rng = np.random.default_rng(3)
    n = 60
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = A + A.conj().T
    v0 = dmrg.start_vector(n, 7)
    E, psi, it = dmrg.lanczos(lambda v: A @ v, v0, 1e-12, 100)
    w, V = np.linalg.eigh(A)
    assert abs(E - w[0]) < 1e-9
    assert abs(abs(np.vdot(V[:, 0], psi)) - 1.0) < 1e-6


def test_lanczos_stopping_rule_beta():
    # an eigenvector start stops at beta <= lanzero after one product
    A = np.diag(np.arange(1.0, 9.0)).astype(complex)
    v0 = np.zeros(8, complex)
    v0[0] = 1.0
    E, psi, it = dmrg.lanczos(lambda v: A @ v, v0, 1e-10, 100)
    assert it == 1 and abs(E - 1.0) < 1e-15


def _rho_from_spectrum(blocks):
    rho = []
    for a, vals in enumerate(blocks):
        rho.append((a, np.diag(np.asarray(vals, dtype=complex))))
    return rho


def test_truncation_keeps_kept_max_when_tail_large():
    rho = _rho_from_spectrum([[0.4, 0.1, 0.05], [0.3, 0.1, 0.05]])
    tr = dmrg.truncation(rho, 2, 3, 1e-18)
    assert tr.num_state == 3
    assert tr.kept_per_sector == [2, 1]
    assert abs(tr.err - 0.2) < 1e-15


def test_truncation_small_dim_keeps_all():
    rho = _rho_from_spectrum([[0.5], [0.5]])
    tr = dmrg.truncation(rho, 4, 8, 1e-18)
    assert tr.num_state == 2 and tr.err == 0.0


def test_truncation_stops_when_tail_below_threshold():
    rho = _rho_from_spectrum([[0.6, 1e-20, 0.0], [0.4, 0.0]])
    tr = dmrg.truncation(rho, 1, 4, float(np.float32(1e-18)))
    assert tr.num_state == 2
    assert tr.kept_per_sector == [1, 1]


def test_truncation_tie_break_is_deterministic():
    # equal weights in two sectors: the earlier sector wins the last slot
    rho = _rho_from_spectrum([[0.25, 0.25], [0.25, 0.25]])
    tr = dmrg.truncation(rho, 1, 3, 1e-18)
    assert tr.num_state == 3
    assert tr.kept_per_sector == [2, 1]


def test_entropy_of_uniform_spectrum():
    rho = _rho_from_spectrum([[0.25, 0.25], [0.25, 0.25]])
    tr = dmrg.truncation(rho, 4, 4, 1e-18)
    assert abs(tr.entropy - math.log(4)) < 1e-14


def test_start_vector_is_deterministic_and_normalised():
    a = dmrg.start_vector(1000, 5)
    b = dmrg.start_vector(1000, 5)
    assert np.array_equal(a, b)
    assert abs(np.vdot(a, a).real - 1.0) < 1e-14
    assert np.all(a.real > 0) and np.all(a.imag > 0)


@pytest.mark.parametrize("kind,N", [("hubbard", 6), ("tj", 4)])
def test_untruncated_triangular_flux_equals_ed(kind, N):
    """Triangular 3 x 3 cylinder with flux (complex diagonal bonds, periodic
    wrap): every finite-sweep step of an untruncated run equals the exact
    ground state."""
    m = lat.Model(kind=kind, Lx=3, Ly=3, lattice="triangular", U=8.0, J=0.4, t=1.0, flux=1.0 / 6, n_elec=N)
    e0 = _ed_gs(m)
    recs = driver.run(m, driver.Schedule(kept_max=4000, n_sweeps=1, tol=1e-12))
    sweep = [r for r in recs if r.phase == "sweep1"]
    assert sweep
    for r in sweep:
        assert abs(r.energy - e0) < 1e-10 * abs(e0), (r.sys_len, r.energy, e0)


def test_lanczos_levels_match_eigh_and_real_projection_does_not():
    """lanczos_levels (lanczos_memo_new with levels > 1) returns the lowest
    eigenvalues of a complex Hermitian matrix in order, mutually orthogonal.
    The reference's real-coefficient projection (Gram_Schmidt_Lanczos,
    code/space_info.f90:1222-1279) planted in its place returns E_0 again as
    level 2: i psi_0 survives it."""
    rng = np.random.default_rng(11)
    n = 80
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = A + A.conj().T
    w, _ = np.linalg.eigh(A)
    starts = [dmrg.start_vector(n, 40 + k) for k in range(3)]
    es, vs, its = dmrg.lanczos_levels(lambda v: A @ v, starts, 1e-13, 100)
    assert np.allclose(es, w[:3], atol=1e-9 * abs(w).max())
    for i in range(3):
        for j in range(i):
            assert abs(np.vdot(vs[i], vs[j])) < 1e-9
    # planted mistake: project the start and Lanczos vectors with real
    # coefficients only (u, not i u)
    orig = dmrg._deflate
    try:
        def real_only(v, lower):
            for u in lower:
                v = v - dmrg._rdot(u, v) * u
            return v
        dmrg._deflate = real_only
        es_r, _, _ = dmrg.lanczos_levels(lambda v: A @ v, starts[:2], 1e-13, 100)
    finally:
        dmrg._deflate = orig
    assert abs(es_r[1] - w[0]) < 1e-6 * abs(w).max() and abs(es_r[1] - w[1]) > 1e-3


def test_lanczos_levels_on_a_superblock_equal_ed_singlet_levels():
    """Two levels of the untruncated Hubbard 4 x 2 superblock (flux 0.15,
    N = 6, S = 0) equal the two lowest S = 0 eigenvalues of the Jordan-Wigner
    exact diagonalisation."""
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.15, n_elec=6)
    run = driver.Run(m, driver.Schedule(kept_max=2000, n_sweeps=0, tol=1e-12))
    run.run()   # blocks of the sweeps, built in the whole lattice
    n = m.nsite
    sg = dmrg.StepGeometry(m, 3, 3)
    se = dmrg.enlarge_in_geometry(run.sys_blocks[3], m, "sys", sg)
    ee = dmrg.enlarge_in_geometry(run.env_blocks[3], m, "env", sg)
    H = dmrg.Superblock(se, ee, sg, m, m.n_elec)
    starts = [dmrg.start_vector(H.sb.dim, 5), dmrg.start_vector(H.sb.dim, 6)]
    es, vs, _ = dmrg.lanczos_levels(H.matvec, starts, 1e-13, 100)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    spec = ed.singlet_spectrum(m, list(range(1, n + 1)), bonds, m.n_elec)
    assert abs(es[0] - spec[0]) < 1e-9 * abs(spec[0])
    assert abs(es[1] - spec[1]) < 1e-9 * abs(spec[1])
    assert spec[1] - spec[0] > 1e-3


def test_density_matrix_levels_mixes_the_two_lowest():
    """Get_density_matrix_*_new with two or more levels: (0.1 rho_1 + 1.9
    rho_2) / 2 block by block with the single-precision literals of the
    reference (trace 1 - 1.1e-8); one level gives rho_1."""
    m = lat.Model(kind="tj", Lx=3, Ly=2, J=0.4, t=1.0, flux=0.1, n_elec=4)
    run = driver.Run(m, driver.Schedule(kept_max=2000, n_sweeps=0, tol=1e-12))
    run.warm_up()
    sg = dmrg.StepGeometry(m, 2, 2)
    se = dmrg.enlarge_in_geometry(run.sys_blocks[2], m, "sys", sg)
    ee = dmrg.enlarge_in_geometry(run.env_blocks[2], m, "env", sg)
    sb = dmrg.super_basis(se.basis, ee.basis, m.n_elec)
    a = dmrg.start_vector(sb.dim, 1)
    b = dmrg.start_vector(sb.dim, 2)
    c = dmrg.start_vector(sb.dim, 3)
    a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
    for side in ("sys", "env"):
        one = dmrg.density_matrix_levels(sb, [a], side)
        mix = dmrg.density_matrix_levels(sb, [a, b, c], side)
        ra, rb = dmrg.density_matrix(sb, a, side), dmrg.density_matrix(sb, b, side)
        assert all(np.array_equal(x[1], y[1]) for x, y in zip(one, ra))
        for (s0, m0), (s1, m1), (s2, m2) in zip(mix, ra, rb):
            assert s0 == s1 == s2
            assert np.allclose(m0, 0.05 * m1 + 0.95 * m2, atol=2e-8)
            assert np.array_equal(m0, (float(np.float32(0.1)) * m1 + float(np.float32(1.9)) * m2) / 2)
        tr = sum(np.trace(x[1]).real for x in mix)
        assert abs(tr - (0.1000000014901161 + 1.899999976158142) / 2) < 1e-13 and abs(tr - 1) > 1e-9
