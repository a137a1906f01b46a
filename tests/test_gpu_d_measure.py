"""GPU parity of the site-density measurement (measure_operator_dia with the
operator chain of Get_sys_oper_dia / Get_env_oper_dia): untruncated runs
against exact diagonalisation, truncated runs against the oracle's chain, and
the sum rule sum_i <n_i> = N."""
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
    """warm-up, LR half sweep, RL sweep; then the centre ground state and its
    density profile"""
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    drv = Driver(ctx, kept_max=kept, tol=tol, kept_max_warm=kept, seed=11, warm_tol_factor=1.0,
                 first_tol_factor=1.0)
    drv.warmup()
    drv.sweep(0, n // 2 - 1, n - 2, tol, TRUN, kept, fresh=True, phase=1)
    drv.sweep(1, n - 2, 1, tol, TRUN, kept, phase=1)
    L = n // 2 - 1
    sb = ctx.bind(drv.block(0, L), drv.block(1, n - L - 2), m.n_elec)
    psi = dev(dmrg.start_vector(sb.dim, 5))
    E, it = ctx.lanczos(sb, psi, 1e-13, 400)
    return E, drv.density(L, psi)


@pytest.mark.parametrize("kind,flux", [("hubbard", 0.15), ("tj", 0.2)])
def test_density_untruncated_equals_ed(ctx, kind, flux):
    m = lat.Model(kind=kind, Lx=4, Ly=2, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=6)
    sg = dmrg.StepGeometry(m, m.nsite // 2 - 1, m.nsite - m.nsite // 2 - 1)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    e0, dens = ed.singlet_ground_density(m, list(range(1, m.nsite + 1)), bonds, m.n_elec)
    E, prof = _gpu_profile(ctx, m, 2000, 1e-12)
    assert abs(E - e0) < 1e-10 * abs(e0)
    assert abs(prof.sum() - m.n_elec) < 1e-10
    assert np.abs(prof - dens).max() < 1e-6, (prof, dens)


def test_density_truncated_matches_oracle_chain(ctx):
    """Hubbard 6x2 ladder with flux at m = 24: the GPU chain against the
    oracle's chain on the oracle's own run of the same schedule.  The two runs
    agree to the truncation level (their kept spaces differ only inside
    degenerate multiplets at the cut), so the bar is 1e-5; the sum rule is
    exact."""
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.1, n_elec=10)
    E, prof = _gpu_profile(ctx, m, 24, 1e-11)
    assert abs(prof.sum() - m.n_elec) < 1e-9
    run = driver.Run(m, driver.Schedule(kept_max=24, n_sweeps=0, tol=1e-11, seed=11, warm_tol_factor=1.0,
                                        first_tol_factor=1.0))
    run.warm_up()
    n = m.nsite
    run.left_to_right(n // 2 - 1, n - 2, 1e-11, TRUN, 24, True, "s")
    run.right_to_left(n - 2, 1, 1e-11, TRUN, 24, "s")
    L = n // 2 - 1
    R = n - L - 2
    sg = dmrg.StepGeometry(m, L, R)
    H = dmrg.Superblock(dmrg.enlarge_in_geometry(run.sys_blocks[L], m, "sys", sg),
                        dmrg.enlarge_in_geometry(run.env_blocks[R], m, "env", sg), sg, m, m.n_elec)
    Eo, psi, _ = dmrg.lanczos(H.matvec, dmrg.start_vector(H.sb.dim, 5), 1e-13, 400)
    ref = measure.density_profile(m, run.sys_blocks, run.env_blocks, run.sys_trun, run.env_trun, L, R, psi, H)
    assert abs(E - Eo) < 1e-7 * abs(Eo), (E, Eo)
    assert np.abs(prof - ref).max() < 1e-5, np.abs(prof - ref).max()
