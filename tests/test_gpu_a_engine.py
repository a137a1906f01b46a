"""GPU parity: the grouped complex GEMM engine and H.psi against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import dmrg, driver
from oracle import lattice as lat

from . import npexec
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


def _rand(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


CUTS = {
    "fixed": None,
    "wide-last": dict(target_items=65536, max_groups=32, edge=1, band=4, wide=1),
    "wide-all": dict(target_items=65536, max_groups=32, edge=0, band=4, wide=2),
    "wide-all-split": dict(target_items=64, max_groups=4, edge=1, band=2, wide=2),
    "split-band8": dict(target_items=64, max_groups=64, edge=0, band=8, wide=0),
}


@pytest.mark.parametrize("cut", list(CUTS))
@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 2e-5)])
def test_engine_random_tiles(ctx, precision, tol, cut):
    """Every operand op (N, T, C, conj), ragged M/N/K, K=0 elementwise terms,
    several terms per tile, the accumulate flag; FP64 and FP32 MFMA; every
    way of cutting the list into work items (32- and 48-column tiles, term
    lists split over many waves, shape classes on and off)."""
    from hofstadter_hubbard_tsc_amd import _lib as L
    if CUTS[cut]:
        ctx.set_launch_config(**CUTS[cut])
    try:
        _random_tiles(ctx, precision, tol)
    finally:
        ctx.set_launch_config(0)


def _random_tiles(ctx, precision, tol):
    from hofstadter_hubbard_tsc_amd import _lib as L
    rng = np.random.default_rng(11)
    A = _rand(rng, 40000)
    B = _rand(rng, 40000)
    out0 = _rand(rng, 60000)
    tiles, terms = [], []
    off = 0
    shapes = [(1, 1, 1), (16, 16, 4), (70, 45, 37), (64, 64, 64), (129, 3, 17), (5, 130, 9), (33, 33, 0)]
    for (M, N, K) in shapes:
        for acc in (0, 1):
            t0 = len(terms)
            for aop in range(4):
                for bop in ((0, 1, 2, 3) if K else (0,)):
                    lda = (M if aop in (0, 3) else K) + 3 if K else M + 2
                    ldb = (K if bop in (0, 3) else N) + 1
                    terms.append((rng.normal(), rng.normal(), int(rng.integers(0, 2000)), int(rng.integers(0, 2000)),
                                  lda, ldb, K, 2, 3, aop if K else (aop if aop in (0, 3) else 0), bop))
            tiles.append((off, M + 5, M, N, t0, len(terms), 1, acc))
            off += (M + 5) * N + 7
    T = np.array(tiles, dtype=L.TILE_DTYPE)
    R = np.array(terms, dtype=L.TERM_DTYPE)
    ref = out0.copy()
    npexec.run(T, R, [None, ref, A, B, None, None, None, None])
    dA, dB, dO = dev(A), dev(B), dev(out0)
    bufs = (C.c_void_p * 8)(None, dO.data_ptr(), dA.data_ptr(), dB.data_ptr(), None, None, None, None)
    rc = L.load().hh_run_tasks(ctx.ptr, T.ctypes.data_as(C.c_void_p), len(T), R.ctypes.data_as(C.c_void_p), len(R),
                               bufs, precision)
    assert rc == 0
    got = host(dO)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < tol


def _oracle_run(model, kept):
    run = driver.Run(model, driver.Schedule(kept_max=kept, n_sweeps=0, tol=1e-8))
    run.warm_up()
    n = model.nsite
    run.left_to_right(n // 2 - 1, n - 2, 1e-8, driver.TRUN_ERR, kept, True, "s0")
    run.right_to_left(n - 2, 1, 1e-8, driver.TRUN_ERR, kept, "s0")
    return run


CASES = [("hubbard", 4, 2, 0.15, 6, 40), ("tj", 4, 4, 0.125, 14, 60), ("hubbard", 6, 2, 0.0, 10, 80)]


@pytest.mark.parametrize("kind,Lx,Ly,flux,N,kept", CASES)
def test_model_wave_matches_oracle(ctx, kind, Lx, Ly, flux, N, kept):
    model = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    ctx.set_model(model)
    run = _oracle_run(model, kept)
    nk = 1 if kind == "hubbard" else 3
    n = model.nsite
    for ns in range(1, n - 2):
        sysb, envb = run.sys_blocks[ns], run.env_blocks[n - ns - 2]
        sg = dmrg.StepGeometry(model, sysb.length, envb.length)
        H = dmrg.Superblock(dmrg.enlarge_in_geometry(sysb, model, "sys", sg),
                            dmrg.enlarge_in_geometry(envb, model, "env", sg), sg, model, N)
        ps = upload_block(ctx, sysb, 0, nk)
        pe = upload_block(ctx, envb, 1, nk)
        sb = ctx.bind(ps, pe, N)
        assert sb.dim == H.sb.dim
        rng = np.random.default_rng(ns)
        psi = _rand(rng, H.sb.dim)
        ref = H.matvec(psi)
        din = dev(psi)
        dout = dev(np.zeros_like(psi))
        ctx.model_wave(sb, din, dout)
        got = host(dout)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13, ns


@pytest.mark.parametrize("kind,Lx,Ly,flux,N,kept", CASES[:2])
def test_lanczos_matches_oracle(ctx, kind, Lx, Ly, flux, N, kept):
    model = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    ctx.set_model(model)
    run = _oracle_run(model, kept)
    nk = 1 if kind == "hubbard" else 3
    n = model.nsite
    ns = n // 2 - 1
    sysb, envb = run.sys_blocks[ns], run.env_blocks[n - ns - 2]
    sg = dmrg.StepGeometry(model, sysb.length, envb.length)
    H = dmrg.Superblock(dmrg.enlarge_in_geometry(sysb, model, "sys", sg),
                        dmrg.enlarge_in_geometry(envb, model, "env", sg), sg, model, N)
    v0 = dmrg.start_vector(H.sb.dim, 77)
    E_ref, psi_ref, it_ref = dmrg.lanczos(H.matvec, v0, 1e-10, 100)
    sb = ctx.bind(upload_block(ctx, sysb, 0, nk), upload_block(ctx, envb, 1, nk), N)
    psi = dev(v0)
    E, it = ctx.lanczos(sb, psi, 1e-10, 100)
    assert abs(E - E_ref) < 1e-10 * abs(E_ref)
    got = host(psi)
    assert abs(abs(np.vdot(got, psi_ref)) - 1.0) < 1e-8


def test_lanczos_refuses_non_finite_blocks(ctx):
    """A block holding a NaN (damaged input) stops the Lanczos solve with an
    error instead of returning a meaningless energy."""
    from hofstadter_hubbard_tsc_amd.dmrg import HHError
    kind, Lx, Ly, flux, N, kept = CASES[0]
    model = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    ctx.set_model(model)
    run = _oracle_run(model, kept)
    n = model.nsite
    ns = n // 2 - 1
    sysb, envb = run.sys_blocks[ns], run.env_blocks[n - ns - 2]
    ps = upload_block(ctx, sysb, 0, 1)
    pe = upload_block(ctx, envb, 1, 1)
    arena = ps.download()
    arena[0] = np.nan
    ps.upload(arena)
    sb = ctx.bind(ps, pe, N)
    psi = dev(_rand(np.random.default_rng(0), sb.dim))
    with pytest.raises(HHError, match="non-finite"):
        ctx.lanczos(sb, psi, 1e-10, 50)


@pytest.mark.parametrize("kind,Lx,Ly,flux,N,kept", CASES[:2])
def test_density_matrix_matches_oracle(ctx, kind, Lx, Ly, flux, N, kept):
    """Get_density_matrix_sys_new / _env_new through the C ABI: every block
    equals the oracle's psi psi^dag (sys) or psi^T conj(psi) (env) of the same
    psi, in super-basis order, and the traces sum to |psi|^2."""
    model = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    ctx.set_model(model)
    run = _oracle_run(model, kept)
    nk = 1 if kind == "hubbard" else 3
    n = model.nsite
    ns = n // 2 - 1
    sysb, envb = run.sys_blocks[ns], run.env_blocks[n - ns - 2]
    sg = dmrg.StepGeometry(model, sysb.length, envb.length)
    H = dmrg.Superblock(dmrg.enlarge_in_geometry(sysb, model, "sys", sg),
                        dmrg.enlarge_in_geometry(envb, model, "env", sg), sg, model, N)
    sb = ctx.bind(upload_block(ctx, sysb, 0, nk), upload_block(ctx, envb, 1, nk), N)
    psi = _rand(np.random.default_rng(7), H.sb.dim)
    psi /= np.linalg.norm(psi)
    for side, name in ((0, "sys"), (1, "env")):
        got = ctx.density_matrix(sb, side, dev(psi))
        ref = dmrg.density_matrix(H.sb, psi, name)
        assert [a for a, _ in got] == [a for a, _ in ref]
        for (_, g), (_, r) in zip(got, ref):
            assert g.shape == r.shape
            assert np.abs(g - r).max() < 1e-14 * max(1.0, np.abs(r).max())
        assert abs(sum(np.trace(g).real for _, g in got) - 1.0) < 1e-12


def test_vector_arguments_are_checked(ctx):
    """A device vector of the wrong dtype or shorter than the superblock is
    refused by the binding before any pointer reaches the library, and
    hh_lanczos_status reports the last solve."""
    import torch
    from hofstadter_hubbard_tsc_amd import HHError
    m = lat.Model(kind="hubbard", Lx=4, Ly=2, U=8.0, t=1.0, flux=0.1, n_elec=6)
    ctx.set_model(m)
    run = driver.Run(m, driver.Schedule(kept_max=20, n_sweeps=0, tol=1e-9))
    run.warm_up()
    n = m.nsite
    sysb, envb = run.sys_blocks[n // 2 - 1], run.env_blocks[n - n // 2 - 1]
    sb = ctx.bind(upload_block(ctx, sysb, 0, 1), upload_block(ctx, envb, 1, 1), m.n_elec)
    good = torch.zeros(sb.dim, dtype=torch.complex128, device="cuda")
    with pytest.raises(HHError):
        ctx.model_wave(sb, good[:-1].contiguous(), good.clone())
    with pytest.raises(HHError):
        ctx.model_wave(sb, torch.zeros(sb.dim, dtype=torch.complex64, device="cuda"), good.clone())
    with pytest.raises(HHError):
        ctx.lanczos(sb, torch.zeros(2 * sb.dim, dtype=torch.float64, device="cuda"), 1e-9, 10)
    psi = dev(dmrg.start_vector(sb.dim, 3))
    E, it = ctx.lanczos(sb, psi, 1e-12, 100)
    conv, dritz, res = ctx.lanczos_status()
    assert conv and dritz <= 1e-12 and res < 1e-5, (conv, dritz, res)
    E2, it2 = ctx.lanczos(sb, dev(dmrg.start_vector(sb.dim, 3)), 1e-12, 3)
    conv, dritz, res = ctx.lanczos_status()
    assert it2 == 3 and not conv and res > 0.0
