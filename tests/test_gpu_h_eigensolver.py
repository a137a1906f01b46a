"""GPU tests of the hand-written density-matrix eigensolver (hh_eigh_sectors
solver 0; the per-sector ZHEEV of code/get_den.f90:242) through the C ABI.

Blocks are packed like the truncation's density matrices: density-matrix-like
Hermitian positive semidefinite blocks (eigenvalues spread over many decades),
random Hermitian blocks, exactly degenerate spectra, diagonal blocks (zero
off-diagonals), a zero block, rank-deficient blocks whose kept set reaches
into the null space (up to every vector of a 1200-wide block), sizes
1 .. 2100 (every back-transformation class).
Bars, against numpy.linalg.eigh on the host:
  * eigenvalues: 1e-13 of the block's largest |eigenvalue|;
  * kept vectors: U^H U = I to 1e-12, |A U - U diag(lambda)| <= 1e-12 |A|
    column by column, and the columns left of the kept ones exactly zero;
  * where a kept eigenvalue is separated from every other by 1e-6 |A|, its
    vector equals numpy's up to a phase to 1e-9;
  * the library eigensolver (solver 1, zheevd) on the same blocks gives the
    same eigenvalues to 1e-13 |A|.
Every step of a truncated half sweep replayed from one state with either
solver: where the discarded weight is above rounding level (1e-14), the same
kept count, the truncation error to 1e-6 relative and the next step's energy
to 1e-10.
"""
import numpy as np
import pytest

from oracle import driver
from oracle import lattice as lat

from .gpuhelp import dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a device"
    from hofstadter_hubbard_tsc_amd import Context
    c = Context(0)
    yield c
    c.close()


def _unitary(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _blocks(rng, kinds):
    out = []
    for kind, n, keep in kinds:
        if kind == "rho":       # psi psi^H with singular values over many decades
            m = max(1, n // 2 + 3)
            s = np.exp(-np.linspace(0.0, 18.0, min(n, m))) if n > 1 else np.ones(1)
            q = _unitary(rng, n)[:, :len(s)]
            a = (q * s ** 2) @ q.conj().T
        elif kind == "herm":
            x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
            a = 0.5 * (x + x.conj().T)
        elif kind == "degenerate":   # eigenvalues in groups of three
            lam = np.repeat(rng.uniform(0.0, 1.0, n // 3 + 1), 3)[:n]
            q = _unitary(rng, n)
            a = (q * lam) @ q.conj().T
        elif kind == "diagonal":
            a = np.diag(rng.uniform(0.0, 1.0, n)).astype(np.complex128)
        elif kind == "zero":
            a = np.zeros((n, n), np.complex128)
        elif kind == "rankdef":     # rank n // 3, more kept than the rank
            r = max(1, n // 3)
            x = rng.normal(size=(n, r)) + 1j * rng.normal(size=(n, r))
            a = x @ x.conj().T / n
        else:
            raise ValueError(kind)
        a = 0.5 * (a + a.conj().T)
        out.append((a, keep))
    return out


def _pack(blocks):
    return np.concatenate([a.reshape(-1, order="F") for a, _ in blocks])


def _run(ctx, blocks, solver):
    import torch
    dims = [a.shape[0] for a, _ in blocks]
    keep = [k for _, k in blocks]
    # only the upper triangle is read: spoil the strictly lower part
    packed = []
    for a, _ in blocks:
        b = a.copy()
        b[np.tril_indices(a.shape[0], -1)] = 7.0 - 3.0j
        packed.append(b.reshape(-1, order="F"))
    t = dev(np.concatenate(packed))
    w = ctx.eigh_sectors(t, dims, keep, solver)
    torch.cuda.synchronize()
    return host(t), w.cpu().numpy()


def _check(ctx, blocks, solver=0):
    arena, w = _run(ctx, blocks, solver)
    off, eo = 0, 0
    for a, keep in blocks:
        n = a.shape[0]
        ref_w, ref_v = np.linalg.eigh(a)
        scale = max(np.abs(ref_w).max(), 1e-300)
        got_w = w[eo:eo + n]
        assert np.abs(got_w - ref_w).max() <= 1e-13 * scale + 1e-300, (n, np.abs(got_w - ref_w).max(), scale)
        blk = arena[off:off + n * n].reshape((n, n), order="F")
        U = blk[:, n - keep:]
        if solver == 0:
            assert not np.any(blk[:, :n - keep]), "columns left of the kept ones must be zero"
        if keep:
            lam = got_w[n - keep:]
            assert np.abs(U.conj().T @ U - np.eye(keep)).max() <= 1e-12, n
            res = np.abs(a @ U - U * lam).max()
            assert res <= 1e-12 * scale + 1e-300, (n, res, scale)
            for j in range(keep):
                i = n - keep + j
                gap = min([abs(ref_w[i] - ref_w[x]) for x in range(n) if x != i] or [np.inf])
                if gap >= 1e-6 * scale:
                    ov = abs(np.vdot(ref_v[:, i], U[:, j]))
                    assert abs(ov - 1.0) <= 1e-9, (n, j, ov, gap)
        off += n * n
        eo += n
    return w


@pytest.mark.parametrize("seed", [1, 2])
def test_small_and_medium_blocks(ctx, seed):
    rng = np.random.default_rng(seed)
    kinds = [("rho", 1, 1), ("rho", 2, 1), ("herm", 3, 3), ("rho", 5, 2), ("herm", 17, 9), ("rho", 31, 31),
             ("degenerate", 32, 16), ("herm", 33, 0), ("diagonal", 40, 13), ("zero", 8, 2), ("rho", 64, 20),
             ("rankdef", 90, 50), ("rho", 100, 40), ("degenerate", 129, 64), ("herm", 257, 100)]
    _check(ctx, _blocks(rng, kinds))


def test_large_blocks_every_class(ctx):
    rng = np.random.default_rng(7)
    kinds = [("rho", 700, 200), ("rho", 1100, 300), ("herm", 1030, 60), ("rho", 2100, 500), ("degenerate", 1500, 30),
             ("rankdef", 1200, 1200), ("rho", 600, 600), ("rankdef", 900, 700)]
    _check(ctx, _blocks(rng, kinds))


def test_agrees_with_zheevd(ctx):
    rng = np.random.default_rng(11)
    kinds = [("rho", 50, 10), ("herm", 300, 40), ("rho", 900, 100)]
    blocks = _blocks(rng, kinds)
    w0 = _check(ctx, blocks, 0)
    w1 = _check(ctx, blocks, 1)
    scale = max(np.abs(w1).max(), 1e-300)
    assert np.abs(w0 - w1).max() <= 1e-13 * scale


def test_dmrg_same_with_either_eigensolver(ctx):
    """Every step of a truncated left-to-right half sweep, replayed from one
    state with each eigensolver: the same kept count and truncation error
    (where the discarded weight is above rounding level), and the next
    step's energy, which depends on the truncation, to 1e-10."""
    from hofstadter_hubbard_tsc_amd import Driver
    m = lat.Model(kind="hubbard", Lx=8, Ly=2, U=8.0, t=1.0, flux=0.05, n_elec=14)
    ctx.set_model(m)
    n = m.nsite
    d = Driver(ctx, kept_max=40, tol=1e-11, kept_max_warm=40, seed=3)
    d.warmup()
    bg = n // 2 - 1
    ctx.set_eigensolver(1)
    d.sweep(0, bg, bg + 1, 1e-11, driver.TRUN_ERR, 40, fresh=True, phase=1)
    compared = 0
    for l in range(bg + 1, n - 3):
        out = []
        for solver in (0, 1):
            f = d.fork()
            ctx.set_eigensolver(solver)
            out.append(f.sweep(0, l, l + 2, 1e-11, driver.TRUN_ERR, 40, phase=1))
        (a0, a1), (b0, b1) = out
        assert a0.energy == b0.energy, (l, a0.energy, b0.energy)   # same state, same solve
        if max(abs(a0.trunc_err), abs(b0.trunc_err)) > 1e-14:
            compared += 1
            assert a0.num_state == b0.num_state, (l, a0.num_state, b0.num_state)
            assert abs(a0.trunc_err - b0.trunc_err) <= 1e-6 * abs(b0.trunc_err), (l, a0.trunc_err, b0.trunc_err)
            assert abs(a1.energy - b1.energy) <= 1e-10 * abs(b1.energy), (l, a1.energy, b1.energy)
        ctx.set_eigensolver(1)
        d.sweep(0, l, l + 1, 1e-11, driver.TRUN_ERR, 40, phase=1)
    ctx.set_eigensolver(0)
    _, handbacks, steps, err = ctx.eigensolver_status()
    print("steps compared %d, hand-backs %d, last orthonormalisation steps %d, |U^H U - I| %.1e"
          % (compared, handbacks, steps, err))
    assert handbacks == 0
    assert compared >= 3
