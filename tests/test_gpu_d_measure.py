"""GPU parity of the site-density measurement (measure_operator_dia with the
operator chain of Get_sys_oper_dia / Get_env_oper_dia) and of the density
correlations <n_i n_j> (Get_density_cor: the pairs across the superblock
centre, and whole rows with the two-site products carried in the system
block), and of the spin correlations <S_i . S_j> and one-body correlations
sum_sigma <c^dag_i c_j> and singlet pair correlations <Delta^dag_a Delta_b>
across the centre, and the every-pair matrices of Get_oper_cor_ndia1:
untruncated runs against exact diagonalisation, truncated
runs against the oracle's chain, and the sum rule sum_i <n_i> = N."""
import numpy as np
import pytest

from oracle import dmrg, driver, ed, measure
from oracle import lattice as lat

from .gpuhelp import dev

pytestmark = pytest.mark.gpu

TRUN = driver.TRUN_ERR


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a device"
    from hofstadter_hubbard_tsc_amd import Context
    c = Context(0)
    yield c
    c.close()


def _gpu_profile(ctx, m, kept, tol):
    """warm-up, LR half sweep, RL sweep; then the centre ground state, its
    density profile and its cross-centre correlations"""
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    drv = Driver(ctx, kept_max=kept, tol=tol, kept_max_warm=kept, seed=11, warm_tol_factor=1.0,
                 first_tol_factor=1.0)
    drv.warmup()
    drv.sweep(0, n // 2 - 1, n - 2, tol, TRUN, kept, fresh=True, phase=1)
    drv.reset_transform()
    drv.sweep(1, n - 2, 1, tol, TRUN, kept, phase=1)
    L = n // 2 - 1
    sb = ctx.bind(drv.block(0, L), drv.block(1, n - L - 2), m.n_elec)
    psi = dev(dmrg.start_vector(sb.dim, 5))
    E, it = ctx.lanczos(sb, psi, 1e-13, 400)
    rows = {r: drv.density_row(L, r, psi) for r in (1, 2, L + 1)}
    srows = {r: drv.spin_row(L, r, psi) for r in (1, 2, L + 1)}
    hrows = {r: drv.hopping_row(L, r, psi) for r in (1, 3, L + 1)}
    return (E, drv.density(L, psi), drv.density_cross(L, psi), rows, drv.spin_cross(L, psi), drv.hopping_cross(L, psi),
            srows, hrows)


@pytest.mark.parametrize("kind,flux", [("hubbard", 0.15), ("tj", 0.2)])
def test_density_untruncated_equals_ed(ctx, kind, flux):
    m = lat.Model(kind=kind, Lx=4, Ly=2, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=6)
    sg = dmrg.StepGeometry(m, m.nsite // 2 - 1, m.nsite - m.nsite // 2 - 1)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    e0, dens, nn = ed.singlet_ground_correlations(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    _, ss = ed.singlet_ground_spin_correlations(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    _, G = ed.singlet_ground_hopping(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    E, prof, cc, rows, sc, hc, srows, hrows = _gpu_profile(ctx, m, 2000, 1e-12)
    assert abs(E - e0) < 1e-10 * abs(e0)
    assert abs(prof.sum() - m.n_elec) < 1e-10
    assert np.abs(prof - dens).max() < 1e-6, (prof, dens)
    left, right = np.arange(m.nsite // 2), np.arange(m.nsite // 2, m.nsite)
    x = np.ix_(left, right)
    assert np.abs(cc[x] - nn[x]).max() < 1e-6, (cc[x], nn[x])
    assert np.array_equal(cc[x], cc.T[x])
    assert np.isnan(cc[np.ix_(left, left)]).all() and np.isnan(cc[np.ix_(right, right)]).all()
    assert np.abs(sc[x] - ss[x]).max() < 1e-6, (sc[x], ss[x])
    assert np.array_equal(sc[x], sc.T[x]) and np.isnan(sc[np.ix_(left, left)]).all()
    assert np.abs(hc[x] - G[x]).max() < 1e-6, (hc[x], G[x])
    assert np.array_equal(hc.T[x], np.conj(hc[x]))
    for r, row in srows.items():   # spin rows: carried products, within-block term, env side
        others = [j for j in range(m.nsite) if j != r - 1]
        assert np.isnan(row[r - 1])
        assert np.abs(row[others] - ss[r - 1, others]).max() < 1e-6, (r, row, ss[r - 1])
    for r, row in hrows.items():   # one-body rows: carried fermionic products, within-block hop, env side
        others = [j for j in range(m.nsite) if j != r - 1]
        assert np.isnan(row[r - 1])
        assert np.abs(row[others] - G[r - 1, others]).max() < 1e-6, (r, row, G[r - 1])
    for r, row in rows.items():   # block-block (carried products), block-centre, env side
        others = [j for j in range(m.nsite) if j != r - 1]
        assert np.isnan(row[r - 1])
        assert np.abs(row[others] - nn[r - 1, others]).max() < 1e-6, (r, row, nn[r - 1])


def test_density_truncated_matches_oracle_chain(ctx):
    """Hubbard 6x2 ladder with flux at m = 24 (truncated): the GPU chain
    against the oracle's chain fed with the GPU run's own blocks, truncation
    matrices and ground state (downloaded), so both sides see identical
    inputs; the sum rule is exact."""
    from .gpuhelp import host, oracle_block, oracle_superblock
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.1, n_elec=10)
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    drv = Driver(ctx, kept_max=24, tol=1e-11, kept_max_warm=24, seed=11, warm_tol_factor=1.0, first_tol_factor=1.0)
    drv.warmup()
    drv.sweep(0, n // 2 - 1, n - 2, 1e-11, TRUN, 24, fresh=True, phase=1)
    drv.reset_transform()
    drv.sweep(1, n - 2, 1, 1e-11, TRUN, 24, phase=1)
    L = n // 2 - 1
    R = n - L - 2
    sb = ctx.bind(drv.block(0, L), drv.block(1, R), m.n_elec)
    psi = dev(dmrg.start_vector(sb.dim, 5))
    ctx.lanczos(sb, psi, 1e-13, 400)
    prof = drv.density(L, psi)
    cc = drv.density_cross(L, psi)
    sc = drv.spin_cross(L, psi)
    hc = drv.hopping_cross(L, psi)
    rows = {r: drv.density_row(L, r, psi) for r in (1, 3, L + 1)}
    srows = {r: drv.spin_row(L, r, psi) for r in (2, L, L + 1)}
    hrows = {r: drv.hopping_row(L, r, psi) for r in (1, L - 1, L + 1)}
    assert abs(prof.sum() - m.n_elec) < 1e-9
    sys_blocks = {l: oracle_block(ctx, drv.block(0, l), m) for l in range(1, L + 1)}
    env_blocks = {l: oracle_block(ctx, drv.block(1, l), m) for l in range(1, R + 1)}
    sys_truns = {l: drv.trunc(0, l).matrices() for l in range(2, L + 1)}
    env_truns = {l: drv.trunc(1, l).matrices() for l in range(2, R + 1)}
    assert any(sum(u.shape[1] for _, u in t) < sum(u.shape[0] for _, u in t) for t in sys_truns.values()), \
        "the case must truncate"
    H = oracle_superblock(ctx, drv.block(0, L), drv.block(1, R), m, m.n_elec)
    ref = measure.density_profile(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H)
    assert np.abs(prof - ref).max() < 1e-10, np.abs(prof - ref).max()
    cref = measure.cross_density(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H)
    fin = ~np.isnan(cref)
    assert np.array_equal(fin, ~np.isnan(cc))
    assert np.abs(cc[fin] - cref[fin]).max() < 1e-10, np.abs(cc[fin] - cref[fin]).max()
    sref = measure.spin_cross(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H)
    fin = ~np.isnan(sref)
    assert np.array_equal(fin, ~np.isnan(sc))
    assert np.abs(sc[fin] - sref[fin]).max() < 1e-10, np.abs(sc[fin] - sref[fin]).max()
    href = measure.hopping_cross(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H)
    fin = ~np.isnan(href)
    assert np.array_equal(fin, ~np.isnan(hc))
    assert np.abs(hc[fin] - href[fin]).max() < 1e-10, np.abs(hc[fin] - href[fin]).max()
    for r, row in srows.items():
        rref = measure.spin_row(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H, r)
        ok = ~np.isnan(rref)
        assert np.array_equal(np.isnan(row), np.isnan(rref))
        assert np.abs(row[ok] - rref[ok]).max() < 1e-10, (r, np.abs(row[ok] - rref[ok]).max())
    for r, row in hrows.items():
        rref = measure.hopping_row(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H, r)
        ok = ~np.isnan(rref)
        assert np.array_equal(np.isnan(row), np.isnan(rref))
        assert np.abs(row[ok] - rref[ok]).max() < 1e-10, (r, np.abs(row[ok] - rref[ok]).max())
    for r, row in rows.items():
        rref = measure.density_row(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H, r)
        assert np.array_equal(np.isnan(row), np.isnan(rref))
        ok = ~np.isnan(rref)
        assert np.abs(row[ok] - rref[ok]).max() < 1e-10, (r, np.abs(row[ok] - rref[ok]).max())


def _pair_bonds(m, L, R):
    sys_b = [(a, b) for a in range(1, L + 1) for b in range(a + 1, L + 1)]
    env_b = [(a, b) for a in range(1, R + 1) for b in range(a + 1, R + 1)]
    return sys_b, env_b, [(m.nsite + 1 - b, m.nsite + 1 - a) for (a, b) in env_b]


def _sweep(ctx, m, kept, tol, L=None):
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    drv = Driver(ctx, kept_max=kept, tol=tol, kept_max_warm=kept, seed=11, warm_tol_factor=1.0, first_tol_factor=1.0)
    drv.warmup()
    drv.sweep(0, n // 2 - 1, n - 2, tol, TRUN, kept, fresh=True, phase=1)
    drv.reset_transform()
    drv.sweep(1, n - 2, 1, tol, TRUN, kept, phase=1)
    L = n // 2 - 1 if L is None else L
    R = n - L - 2
    sb = ctx.bind(drv.block(0, L), drv.block(1, R), m.n_elec)
    psi = dev(dmrg.start_vector(sb.dim, 5))
    E, _ = ctx.lanczos(sb, psi, 1e-13, 400)
    return drv, L, R, E, psi


@pytest.mark.parametrize("kind,flux,Lx,N", [("hubbard", 0.15, 4, 6), ("tj", 0.2, 4, 6)])
def test_pair_cross_untruncated_equals_ed(ctx, kind, flux, Lx, N):
    """<Delta^dag_a Delta_b> (Get_mSCOP_Operator_cor0's block-block pairs) for
    every site pair a of the sys block and b of the env block against exact
    diagonalisation; bonds on one side only come back NaN."""
    m = lat.Model(kind=kind, Lx=Lx, Ly=2, U=4.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    drv, L, R, E, psi = _sweep(ctx, m, 2000, 1e-12)
    sg = dmrg.StepGeometry(m, L, R)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    sys_b, env_b, env_pos = _pair_bonds(m, L, R)
    e0, D = ed.singlet_ground_pairing(m, list(range(1, m.nsite + 1)), bonds, m.n_elec, sys_b, env_pos)
    assert abs(E - e0) < 1e-10 * abs(e0)
    allb = sys_b + env_pos
    P = drv.pair_cross(L, psi, allb)
    ns = len(sys_b)
    got = P[:ns, ns:]
    assert np.abs(got - D).max() < 1e-6, (got, D)
    assert np.array_equal(P[ns:, :ns], np.conj(got).T)
    assert np.isnan(P[:ns, :ns]).all() and np.isnan(P[ns:, ns:]).all()


def test_pair_cross_truncated_matches_oracle(ctx):
    """Hubbard 6 x 2 ladder with flux at m = 24: the GPU pair chains and
    correlations against the oracle's on the GPU run's own blocks,
    truncations and ground state."""
    from .gpuhelp import host, oracle_block, oracle_superblock
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=4.0, t=1.0, flux=0.1, n_elec=10)
    drv, L, R, E, psi = _sweep(ctx, m, 24, 1e-11)
    sys_b, env_b, env_pos = _pair_bonds(m, L, R)
    sys_b, env_b, env_pos = sys_b[::3], env_b[::3], env_pos[::3]
    P = drv.pair_cross(L, psi, sys_b + env_pos)[:len(sys_b), len(sys_b):]
    sys_blocks = {l: oracle_block(ctx, drv.block(0, l), m) for l in range(1, L + 1)}
    env_blocks = {l: oracle_block(ctx, drv.block(1, l), m) for l in range(1, R + 1)}
    sys_truns = {l: drv.trunc(0, l).matrices() for l in range(2, L + 1)}
    env_truns = {l: drv.trunc(1, l).matrices() for l in range(2, R + 1)}
    H = oracle_superblock(ctx, drv.block(0, L), drv.block(1, R), m, m.n_elec)
    ref = measure.pair_cross(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H, sys_b, env_b)
    assert np.abs(ref).max() > 1e-4
    assert np.abs(P - ref).max() < 1e-10, np.abs(P - ref).max()


def test_pair_same_block_untruncated_equals_ed(ctx):
    """<Delta^dag_a Delta_b> for bonds a, b inside the sys block (Get_sys_two1's
    four-site operators) against exact diagonalisation: t-J 6 x 2, 4 electrons."""
    m = lat.Model(kind="tj", Lx=6, Ly=2, J=0.4, t=1.0, flux=0.2, n_elec=4)
    drv, L, R, E, psi = _sweep(ctx, m, 4000, 1e-12)
    sg = dmrg.StepGeometry(m, L, R)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    quads = [((a, b), (c, d)) for a in range(1, L + 1) for b in range(a + 1, L + 1)
             for c in range(b + 1, L + 1) for d in range(c + 1, L + 1)]
    e0, D = ed.singlet_ground_pairing(m, list(range(1, m.nsite + 1)), bonds, m.n_elec,
                                      [q[0] for q in quads], [q[1] for q in quads])
    assert abs(E - e0) < 1e-10 * abs(e0)
    got = drv.pair_block(L, psi, quads)
    assert np.abs(got - np.diag(D)).max() < 1e-6, (got, np.diag(D))


def test_pair_same_block_truncated_matches_oracle(ctx):
    """Hubbard 6 x 2 ladder with flux at m = 24: the GPU four-site chains against
    the oracle's on the GPU run's own blocks, truncations and ground state."""
    from .gpuhelp import host, oracle_block, oracle_superblock
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=4.0, t=1.0, flux=0.1, n_elec=10)
    drv, L, R, E, psi = _sweep(ctx, m, 24, 1e-11)
    quads = [((a, b), (c, d)) for a in range(1, L + 1) for b in range(a + 1, L + 1)
             for c in range(b + 1, L + 1) for d in range(c + 1, L + 1)]
    got = drv.pair_block(L, psi, quads)
    sys_blocks = {l: oracle_block(ctx, drv.block(0, l), m) for l in range(1, L + 1)}
    sys_truns = {l: drv.trunc(0, l).matrices() for l in range(2, L + 1)}
    H = oracle_superblock(ctx, drv.block(0, L), drv.block(1, R), m, m.n_elec)
    ref = measure.pair_same_block(m, sys_blocks, sys_truns, L, host(psi), H, quads)
    assert np.abs(ref).max() > 1e-4
    assert np.abs(got - ref).max() < 1e-10, np.abs(got - ref).max()


@pytest.mark.parametrize("kind,flux,Lx,cut", [("hubbard", 0.15, 4, None), ("hubbard", 0.15, 4, 2),
                                              ("tj", 0.2, 5, None), ("tj", 0.2, 5, 5), ("hubbard", 0.15, 4, 1),
                                              ("tj", 0.2, 5, 7)])
def test_all_pairs_untruncated_equal_ed(ctx, kind, flux, Lx, cut):
    """Get_oper_cor_ndia1 for the density, the spin and the one-body
    correlation (every pair and the on-site values) against exact
    diagonalisation at symmetric and asymmetric cuts; symmetry, Hermiticity
    and the sum rules sum_j <n_i n_j> = N <n_i>, sum_j <S_i . S_j> = 0."""
    m = lat.Model(kind=kind, Lx=Lx, Ly=2, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=6)
    drv, L, R, E, psi = _sweep(ctx, m, 4000, 1e-12, cut)
    sg = dmrg.StepGeometry(m, L, R)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    pos = list(range(1, m.nsite + 1))
    e0, dens, nn = ed.singlet_ground_correlations(m, pos, bonds, m.n_elec)
    _, ss = ed.singlet_ground_spin_correlations(m, pos, bonds, m.n_elec)
    _, G = ed.singlet_ground_hopping(m, pos, bonds, m.n_elec)
    assert abs(E - e0) < 1e-10 * abs(e0)
    A, B, C = drv.density_all(L, psi), drv.spin_all(L, psi), drv.hopping_all(L, psi)
    assert not np.isnan(A).any() and not np.isnan(B).any() and not np.isnan(C).any()
    assert np.abs(A - nn).max() < 1e-6, np.abs(A - nn).max()
    assert np.abs(B - ss).max() < 1e-6, np.abs(B - ss).max()
    assert np.abs(C - G).max() < 1e-6, np.abs(C - G).max()
    assert np.array_equal(A, A.T) and np.array_equal(B, B.T) and np.array_equal(C, np.conj(C.T))
    assert np.abs(A.sum(axis=1) - m.n_elec * np.diag(C).real).max() < 1e-9
    assert np.abs(B.sum(axis=1)).max() < 1e-9


@pytest.mark.parametrize("cut", [None, 3])
def test_all_pairs_truncated_match_oracle(ctx, cut):
    """Hubbard 6 x 2 ladder with flux at m = 24 (truncated): every entry of
    the three correlation matrices against the oracle's all-pairs chains fed
    with the GPU run's own blocks, truncations and ground state."""
    from .gpuhelp import host, oracle_block, oracle_superblock
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=4.0, t=1.0, flux=0.1, n_elec=10)
    drv, L, R, E, psi = _sweep(ctx, m, 24, 1e-11, cut)
    got = {"n": drv.density_all(L, psi), "S": drv.spin_all(L, psi), "cdag": drv.hopping_all(L, psi)}
    sys_blocks = {l: oracle_block(ctx, drv.block(0, l), m) for l in range(1, L + 1)}
    env_blocks = {l: oracle_block(ctx, drv.block(1, l), m) for l in range(1, R + 1)}
    sys_truns = {l: drv.trunc(0, l).matrices() for l in range(2, L + 1)}
    env_truns = {l: drv.trunc(1, l).matrices() for l in range(2, R + 1)}
    assert any(sum(u.shape[1] for _, u in t) < sum(u.shape[0] for _, u in t)
               for t in list(sys_truns.values()) + list(env_truns.values())), "the case must truncate"
    H = oracle_superblock(ctx, drv.block(0, L), drv.block(1, R), m, m.n_elec)
    for name, g in got.items():
        ref = measure.corr_all(m, sys_blocks, env_blocks, sys_truns, env_truns, L, R, host(psi), H, name)
        assert np.abs(g - ref).max() < 1e-10, (name, np.abs(g - ref).max())
