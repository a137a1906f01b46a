"""GPU parity of multi-level targeting (levels > 1) through the C ABI:
lanczos_memo_new with several levels (code/lanczos.f90:29-44,
Gram_Schmidt_Lanczos code/space_info.f90:1222-1279) and the mixed density
matrix of two levels (Get_density_matrix_sys_new / _env_new,
code/get_den.f90:103-177) with its truncation, against the oracle on
identical blocks and against exact diagonalisation; and the sweep driver
targeting two levels in every step (Get_Ground_State with levels,
code/quantum.f90:179-288), with seeded or carried starts of level 2."""
import numpy as np
import pytest

from oracle import dmrg, driver, ed
from oracle import lattice as lat

from .gpuhelp import dev, host, upload_block

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a device"
    from hofstadter_hubbard_tsc_amd import Context
    c = Context(0)
    yield c
    c.close()


def _superblock(ctx, model, kept, ns):
    """Oracle blocks of a finished sweep (built in the whole lattice), the
    oracle superblock of sys block ns and its GPU twin."""
    run = driver.Run(model, driver.Schedule(kept_max=kept, n_sweeps=0, tol=1e-12))
    run.run()
    n = model.nsite
    sysb, envb = run.sys_blocks[ns], run.env_blocks[n - ns - 2]
    sg = dmrg.StepGeometry(model, sysb.length, envb.length)
    H = dmrg.Superblock(dmrg.enlarge_in_geometry(sysb, model, "sys", sg),
                        dmrg.enlarge_in_geometry(envb, model, "env", sg), sg, model, model.n_elec)
    nk = 1 if model.kind == "hubbard" else 3
    ctx.set_model(model)
    sb = ctx.bind(upload_block(ctx, sysb, 0, nk), upload_block(ctx, envb, 1, nk), model.n_elec)
    assert sb.dim == H.sb.dim
    return H, sb, sg


def test_levels_equal_ed_singlet_levels_and_the_oracle(ctx):
    """Three levels of the untruncated Hubbard 4 x 2 superblock (flux 0.15,
    N = 6, S = 0): the energies are the three lowest S = 0 eigenvalues of the
    Jordan-Wigner exact diagonalisation (1e-10), equal the oracle's
    lanczos_levels from the same start vectors, and the vectors are
    orthonormal in the complex inner product (1e-10)."""
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.15, n_elec=6)
    H, sb, sg = _superblock(ctx, m, 2000, 3)
    starts = [dmrg.start_vector(H.sb.dim, 90 + k) for k in range(3)]
    es_o, vs_o, _ = dmrg.lanczos_levels(H.matvec, starts, 1e-13, 100)
    psis = [dev(v) for v in starts]
    es, its = ctx.lanczos_levels(sb, psis, 1e-13, 100)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    spec = ed.singlet_spectrum(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    print("levels", es, "ED", spec[:3], "Lanczos vectors", its)
    for k in range(3):
        assert abs(es[k] - spec[k]) < 1e-10 * abs(spec[k]), (k, es[k], spec[k])
        assert abs(es[k] - es_o[k]) < 1e-10 * abs(es_o[k])
    V = [host(p) for p in psis]
    for i in range(3):
        assert abs(np.vdot(V[i], V[i]) - 1.0) < 1e-10
        for j in range(i):
            assert abs(np.vdot(V[i], V[j])) < 1e-10, (i, j)
        # each vector is the oracle's up to a phase (levels are not degenerate)
        assert abs(abs(np.vdot(vs_o[i], V[i])) - 1.0) < 1e-8


def test_mixed_density_matrix_and_truncation_match_the_oracle(ctx):
    """Two levels of a truncated t-J 4 x 4 cylinder superblock (flux 0.125):
    the mixed density matrix (0.1 rho_1 + 1.9 rho_2) / 2 equals the oracle's
    block by block (1e-14 of its largest element), on both sides, and the
    truncation on it keeps the oracle's kept count and per-sector counts with
    its truncation error (1e-3 relative)."""
    m = lat.Model(kind="tj", Lx=4, Ly=4, J=0.4, t=1.0, flux=0.125, n_elec=14)
    H, sb, _ = _superblock(ctx, m, 40, 7)
    starts = [dmrg.start_vector(H.sb.dim, 5), dmrg.start_vector(H.sb.dim, 6)]
    psis = [dev(v) for v in starts]
    ctx.lanczos_levels(sb, psis, 1e-12, 100)
    p1, p2 = host(psis[0]), host(psis[1])
    for side, name in ((0, "sys"), (1, "env")):
        got = ctx.density_matrix(sb, side, psis[0], psis[1])
        ref = dmrg.density_matrix_levels(H.sb, [p1, p2], name)
        assert len(got) == len(ref)
        scale = max(np.abs(r).max() for _, r in ref)
        for (a, g), (b, r) in zip(got, ref):
            assert a == b
            assert np.abs(g - r).max() <= 1e-14 * scale
        kmin, kmax, terr = 30, 40, 1e-8
        t = ctx.truncate(sb, side, psis[0], kmin, kmax, terr, psi2=psis[1])
        tr = dmrg.truncation(ref, kmin, kmax, terr)
        assert t.num_state == tr.num_state
        assert abs(t.err - tr.err) <= 1e-3 * max(tr.err, 1e-300)
        # per-sector kept counts, wherever the cut falls in a gap of the
        # spectrum (a tie across sectors may be split either way)
        ev = np.sort(np.concatenate([np.linalg.eigvalsh(0.5 * (r + r.conj().T)) for _, r in ref]))[::-1]
        if tr.num_state < len(ev) and ev[tr.num_state - 1] - ev[tr.num_state] > 1e-9 * ev[0]:
            kept_gpu = {int(a): int(k) for a, k in zip(*t.kept()) if k > 0}
            kept_ref = {a: c for (a, _), c in zip(ref, tr.kept_per_sector) if c > 0}
            assert kept_gpu == kept_ref


def test_driver_levels_sweep_equals_ed_singlet_levels(ctx):
    """The step-wise driver with levels = 2 (Get_Ground_State with levels,
    code/quantum.f90:179-288): every step of the first full left-to-right
    sweep over the untruncated Hubbard 4 x 2 lattice (flux 0.15, N = 6) reports the two lowest
    S = 0 levels of the exact diagonalisation (1e-9), the record's energy is
    level 1's, and the levels stay set on a fork; several levels refuse
    solvers 1 and 2."""
    from hofstadter_hubbard_tsc_amd import Driver, HHError
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.15, n_elec=6)
    _, _, sg = _superblock(ctx, m, 2000, 3)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    spec = ed.singlet_spectrum(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    drv = Driver(ctx, kept_max=2000, tol=1e-12, kept_max_warm=2000, seed=11)
    drv.set_levels(2)
    w = drv.warmup()
    assert drv.level_energies().shape == (len(w), 2)
    n = m.nsite
    # the warm-up's blocks hold the warm-up's own geometry: the half sweeps
    # out from the centre and back rebuild every block in the lattice, after
    # which each step's superblock is the whole untruncated lattice
    drv.sweep(0, n // 2 - 1, n - 2, 1e-12, driver.TRUN_ERR, 2000, fresh=True, phase=1)
    drv.reset_transform()
    drv.sweep(1, n - 2, 1, 1e-12, driver.TRUN_ERR, 2000, phase=1)
    drv.reset_transform()
    recs = drv.sweep(0, 1, n - 2, 1e-12, driver.TRUN_ERR, 2000, phase=2)
    E = drv.level_energies()
    assert E.shape == (len(recs), 2) and len(recs) == n - 3
    for r, (e1, e2) in zip(recs, E):
        print("sys_len=%d dim=%d E1=%.12f E2=%.12f (ED %.12f %.12f)" % (r.sys_len, r.dim, e1, e2, spec[0], spec[1]))
        assert r.energy == e1
        # the record's H.psi count covers both levels, its iterations level 1's
        assert r.hpsi_fp64 > r.iters
        assert abs(e1 - spec[0]) < 1e-9 * abs(spec[0])
        assert abs(e2 - spec[1]) < 1e-9 * abs(spec[1])
    f = drv.fork()
    assert f.levels == 2
    f.reset_transform()
    f.sweep(1, n - 2, n - 3, 1e-12, driver.TRUN_ERR, 2000, phase=1)
    assert abs(f.level_energies()[0, 1] - spec[1]) < 1e-9 * abs(spec[1])
    with pytest.raises(HHError):
        drv.set_solver(2)
    drv.set_levels(1)
    drv.set_solver(2)
    with pytest.raises(HHError):
        drv.set_levels(2)


def test_driver_carried_level_starts_give_the_same_levels_with_fewer_hpsi(ctx):
    """Levels 2.. started from the wave-transformed previous levels
    (set_level_start(True)) instead of seeded vectors: on the untruncated
    Hubbard 4 x 2 schedule every step of the full sweep still reports the two
    lowest ED singlet levels (1e-9), with fewer H.psi over the sweep."""
    from hofstadter_hubbard_tsc_amd import Driver
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.15, n_elec=6)
    _, _, sg = _superblock(ctx, m, 2000, 3)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    spec = ed.singlet_spectrum(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    n = m.nsite
    runs = {}
    for carry in (False, True):
        drv = Driver(ctx, kept_max=2000, tol=1e-12, kept_max_warm=2000, seed=11)
        drv.set_levels(2)
        drv.set_level_start(carry)
        drv.warmup()
        drv.sweep(0, n // 2 - 1, n - 2, 1e-12, driver.TRUN_ERR, 2000, fresh=True, phase=1)
        drv.reset_transform()
        drv.sweep(1, n - 2, 1, 1e-12, driver.TRUN_ERR, 2000, phase=1)
        drv.reset_transform()
        recs = drv.sweep(0, 1, n - 2, 1e-12, driver.TRUN_ERR, 2000, phase=2)
        E = drv.level_energies()
        for e1, e2 in E:
            assert abs(e1 - spec[0]) < 1e-9 * abs(spec[0]) and abs(e2 - spec[1]) < 1e-9 * abs(spec[1])
        runs[carry] = sum(r.hpsi_fp64 for r in recs)
    print("H.psi over the sweep: seeded %d, carried %d" % (runs[False], runs[True]))
    assert runs[True] < runs[False]


def test_driver_levels_truncate_on_the_mixed_density_matrix(ctx):
    """A truncated t-J 4 x 4 cylinder sweep (flux 0.125, m = 40) with levels
    = 2: each step's kept count, truncation error and entropy are those of the
    mixed density matrix (0.1 rho_1 + 1.9 rho_2) / 2 of the step's superblock,
    recomputed from that superblock's two lowest levels (kept count exact,
    error 1e-4 relative, entropy 1e-6), and not those of rho_1 alone."""
    from hofstadter_hubbard_tsc_amd import Driver
    m = lat.Model(kind="tj", Lx=4, Ly=4, J=0.4, t=1.0, flux=0.125, n_elec=14)
    ctx.set_model(m)
    kmax = 40
    drv = Driver(ctx, kept_max=kmax, tol=1e-12, kept_max_warm=kmax, seed=5)
    drv.set_levels(2)
    drv.warmup()
    n = m.nsite
    checked = 0
    for ls in range(n // 2 - 1, n // 2 + 2):
        r = drv.sweep(0, ls, ls + 1, 1e-12, driver.TRUN_ERR, kmax, fresh=(ls == n // 2 - 1), phase=1)[-1]
        e1, e2 = drv.level_energies()[-1]
        sl, el, psi = drv.wave()
        sb = ctx.bind(drv.block(0, sl), drv.block(1, el), m.n_elec)
        psis = [psi.clone(), dev(dmrg.start_vector(sb.dim, 40 + ls))]
        es, _ = ctx.lanczos_levels(sb, psis, 1e-13, 100)
        print("sys_len=%d dim=%d driver E=(%.12f, %.12f) again (%.12f, %.12f) kept=%d err=%.6e S=%.9f"
              % (sl, r.dim, e1, e2, es[0], es[1], r.num_state, r.trunc_err, r.entropy))
        assert abs(es[0] - e1) < 1e-9 * abs(e1) and abs(es[1] - e2) < 1e-9 * abs(e2)
        kmin = min(3 * kmax // 4, 3600)
        t = ctx.truncate(sb, 0, psis[0], kmin, kmax, driver.TRUN_ERR, psi2=psis[1])
        t1 = ctx.truncate(sb, 0, psis[0], kmin, kmax, driver.TRUN_ERR)
        assert t.num_state == r.num_state
        assert abs(t.err - r.trunc_err) <= 1e-4 * r.trunc_err
        assert abs(t.entropy - r.entropy) <= 1e-6
        assert abs(t1.err - r.trunc_err) > 1e-2 * r.trunc_err
        checked += 1
    assert checked == 3
