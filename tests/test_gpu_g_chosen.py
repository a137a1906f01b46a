"""GPU tests of the chosen form of the ground-state solve, through the C ABI.

* FP32-storage H.psi (complex64 copies of every operand, FP32 MFMA) against
  the oracle's FP64 H.psi: 1e-5 of max |H psi| (the form's stated bound; it
  measures about 2e-7 at m = 2000).
* The preconditioner's HIP kernels (features, MLP forward, head,
  preconditioner, cosine loss, backward) against the numpy restatement
  (tests/precond_ref.py, itself pinned by finite differences): features 1e-5,
  outputs 1e-4, loss 1e-5 relative, gradient 1e-4 of its largest entry,
  including sectors split over several reduction chunks.
* Solvers 1 and 2 (learned, diagonal and no preconditioner) from one state
  (driver forks): the first step's energy equals the plain FP64 Lanczos's to
  1e-9 relative with all solves converged at tolerance 1e-11; the energy a
  mixed solve reports is the FP64 Rayleigh quotient of the vector it returns
  (1e-12 relative, checked with hh_model_wave in FP64).
* The fallback to the mixed Lanczos, forced from the host by an unreachable
  residual-reduction factor: flagged in the records, energies unchanged.
* A network copied between drivers acts exactly like its source.
* A windowed fork (hh_driver_fork_window) sweeps the steps inside its window
  exactly like a full fork and refuses a step outside it.
* Below the hand-over size (hh_driver_set_min_dim) solver 2 is the plain
  Lanczos, bit for bit.
* Training: the loss falls; what the trained network gains over the method
  without it (total H.psi over the same steps) is printed and bounded.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import driver
from oracle import lattice as lat

from . import precond_ref as R
from .gpuhelp import host, oracle_superblock

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


def _prepared(ctx, m, kept, capture=0):
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    d = Driver(ctx, kept_max=kept, tol=1e-11, kept_max_warm=kept, seed=17)
    d.set_min_dim(0)   # these superblocks are far below the default hand-over size
    d.warmup()
    bg = n // 2 - 1
    if capture:
        d.capture(capture)
        d.set_solver(2)
        d.set_precond(1)
    d.sweep(0, bg, bg + 2, 1e-11, TRUN, kept, fresh=True, phase=1)
    d.set_solver(0)
    d.set_precond(0)
    return d, bg + 2


@pytest.mark.parametrize("wide", [0, 2])
def test_fp32_storage_matvec_matches_oracle(ctx, wide):
    """FP32-storage H.psi against the oracle, with 32-column tiles and with
    every tile 48 columns wide (and FP64 H.psi of the wide cut to 1e-13)."""
    import torch
    m = lat.Model(kind="tj", Lx=4, Ly=4, J=0.4, t=1.0, flux=0.125, n_elec=14)
    d, _ = _prepared(ctx, m, 40)
    ls, le, psi = d.wave()
    sysb, envb = d.block(0, ls), d.block(1, le)
    if wide:
        ctx.set_launch_config(target_items=4096, max_groups=8, edge=1, band=4, wide=wide)
    try:
        sb = ctx.bind(sysb, envb, m.n_elec)
    finally:
        ctx.set_launch_config(0)
    H = oracle_superblock(ctx, sysb, envb, m, m.n_elec)
    ref = H.matvec(host(psi))
    y = torch.empty_like(psi)
    ctx.model_wave(sb, psi, y)
    assert np.abs(host(y) - ref).max() <= 1e-13 * np.abs(ref).max()
    ctx.set_precision(2)
    try:
        ctx.model_wave(sb, psi, y)
    finally:
        ctx.set_precision(0)
    err = np.abs(host(y) - ref).max() / np.abs(ref).max()
    print("FP32-storage H.psi relative error %.2e (wide %d)" % (err, wide))
    assert err <= 1e-5


def _eval(ctx, prob):
    from hofstadter_hubbard_tsc_amd import _lib as L
    rows, cols, s2, nq, nmid, hs, he, r, delta, theta, w = prob
    lib = L.load()
    np_ = len(rows)
    rr = np.ascontiguousarray(np.stack([r.real, r.imag], 1).reshape(-1))
    dd = np.ascontiguousarray(np.stack([delta.real, delta.imag], 1).reshape(-1))
    wf = np.ascontiguousarray(w, dtype=np.float32)
    feat = np.zeros(np_ * R.NF, np.float32)
    out = np.zeros(np_ * R.NO, np.float32)
    par = np.zeros(2 * np_)
    loss = np.zeros(1)
    grad = np.zeros(R.SIZE, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    rc = lib.hh_precond_eval(ctx.ptr, np_, L.ip(rows), L.ip(cols), L.ip(s2), L.ip(nq), nmid, L.ip(hs), L.ip(he),
                             L.ip(rr), L.ip(dd), theta, fp(wf), fp(feat), fp(out), L.ip(par), L.ip(loss), fp(grad))
    assert rc == 0, lib.hh_last_error(ctx.ptr)
    return feat.reshape(-1, R.NF), out.reshape(-1, R.NO), par.reshape(-1, 2), loss[0], grad


@pytest.mark.parametrize("seed,big", [(1, False), (2, False), (3, True)])
def test_precond_kernels_match_numpy(ctx, seed, big):
    from hofstadter_hubbard_tsc_amd import _lib as L
    assert L.load().hh_precond_size() == R.SIZE
    prob = list(R.random_problem(seed))
    if big:
        # one sector of 256 x 200 elements spans two reduction chunks
        rng = np.random.default_rng(seed)
        prob[0] = np.array([256, 3, 5], np.int32)
        prob[1] = np.array([200, 4, 2], np.int32)
        prob[2] = np.array([2, 0, 4], np.int32)
        prob[3] = np.array([14, 12, 13], np.int32)
        prob[5] = rng.normal(2.0, 1.0, int(prob[0].sum()))
        prob[6] = rng.normal(1.0, 1.0, int(prob[1].sum()))
        n = int(np.sum(prob[0] * prob[1]))
        prob[7] = rng.normal(size=n) + 1j * rng.normal(size=n)
        prob[8] = prob[7] / (2.0 + np.abs(rng.normal(size=n))) + 0.1 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    prob[10] = np.asarray(prob[10], np.float32).astype(np.float64)
    feat, out, par, loss, grad = _eval(ctx, prob)
    Lr, gr, Fr, outr, sig, al = R.loss_and_grad(*prob)
    assert np.abs(feat - Fr).max() <= 1e-5
    assert np.abs(out - outr).max() <= 1e-4 * max(1.0, np.abs(outr).max())
    assert np.allclose(par[:, 0], sig, rtol=1e-4) and np.allclose(par[:, 1], al, rtol=1e-4)
    assert abs(loss - Lr) <= 1e-5 * abs(Lr)
    assert np.abs(grad - gr).max() <= 1e-4 * np.abs(gr).max()


def test_forms_reach_the_plain_energy_and_report_exact_energies(ctx):
    import torch
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.05, n_elec=10)
    d, start = _prepared(ctx, m, 32, capture=8)
    losses, ns = d.train_precond(epochs=40, lr=1e-3)
    assert ns >= 2 and losses[-1] <= losses[0]
    forms = {}
    from hofstadter_hubbard_tsc_amd.dmrg import HHError
    for name, solver, mode, kmax in (("plain", 0, 0, 0), ("mixed", 1, 0, 0), ("learned", 2, 0, 0),
                                     ("diagonal", 2, 1, 0), ("none", 2, 2, 0), ("diagonal_k8", 2, 1, 8)):
        f = d.fork()
        f.set_solver(solver)
        f.set_precond(mode)
        if kmax:
            # a subspace of 8 restarts the Davidson every seven expansions
            for bad in (1, 129):
                with pytest.raises(HHError):
                    f.set_subspace(bad)
            f.set_subspace(kmax)
        forms[name] = f
    first = {}
    for name, f in forms.items():
        recs = f.sweep(0, start, start + 1, 1e-11, TRUN, 32, phase=2)
        r = recs[0]
        assert r.converged, (name, r.iters)
        first[name] = r.energy
        # the reported energy is the FP64 Rayleigh quotient of the returned vector
        ls, le, psi = f.wave()
        sb = ctx.bind(f.block(0, ls), f.block(1, le), m.n_elec)
        y = torch.empty_like(psi)
        ctx.model_wave(sb, psi, y)
        rq = torch.vdot(psi, y).real.item() / torch.vdot(psi, psi).real.item()
        assert abs(rq - r.energy) <= 1e-12 * abs(r.energy), (name, rq, r.energy)
        print("%s: E = %.13f, H.psi fp64 %d fp32 %d" % (name, r.energy, r.hpsi_fp64, r.hpsi_fp32))
    for name, e in first.items():
        assert abs(e - first["plain"]) <= 1e-9 * abs(first["plain"]), (name, e, first["plain"])


def test_copied_network_acts_like_its_source(ctx):
    """A driver forked before training takes the trained network by
    hh_driver_copy_precond and then sweeps exactly like a fork made after
    training (same energies, same H.psi counts)."""
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.05, n_elec=10)
    d, start = _prepared(ctx, m, 32, capture=8)
    early = d.fork()
    from hofstadter_hubbard_tsc_amd.dmrg import HHError
    with pytest.raises(HHError):
        early.copy_precond(early)      # no network yet
    d.train_precond(epochs=20, lr=1e-3)
    late = d.fork()
    early.copy_precond(d)
    out = []
    for f in (early, late):
        f.set_solver(2)
        f.set_precond(0)
        out.append(f.sweep(0, start, start + 2, 1e-11, TRUN, 32, phase=2))
    for a, b in zip(*out):
        assert a.energy == b.energy and a.hpsi_fp64 == b.hpsi_fp64 and a.hpsi_fp32 == b.hpsi_fp32


def test_small_superblocks_go_to_the_plain_solver(ctx):
    """Below the hand-over size solver 2 runs the plain Lanczos: the same
    energies, iteration counts and H.psi counts as solver 0, bit for bit."""
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.05, n_elec=10)
    d, start = _prepared(ctx, m, 32)
    a, b = d.fork(), d.fork()
    a.set_solver(0)
    b.set_solver(2)
    b.set_min_dim(10 ** 9)
    ra = a.sweep(0, start, start + 2, 1e-11, TRUN, 32, phase=2)
    rb = b.sweep(0, start, start + 2, 1e-11, TRUN, 32, phase=2)
    for x, y in zip(ra, rb):
        assert x.energy == y.energy and x.iters == y.iters and x.hpsi_fp64 == y.hpsi_fp64 and y.hpsi_fp32 == 0


def test_windowed_fork_sweeps_like_a_full_fork(ctx):
    """hh_driver_fork_window: a copy holding only the blocks and truncations
    that the next two left-to-right steps touch sweeps them exactly like a
    full copy (energies, iteration counts, kept counts); a third step, whose
    env block lies outside the window, fails with status 1; a window that
    leaves out the current superblock's blocks is refused."""
    from hofstadter_hubbard_tsc_amd.dmrg import HHError
    m = lat.Model(kind="hubbard", Lx=6, Ly=2, U=8.0, t=1.0, flux=0.05, n_elec=10)
    d, start = _prepared(ctx, m, 32)
    n, steps = m.nsite, 2
    # a step at sys length l reads sys block l, env block n - l - 2, the sys
    # truncation of l and the env truncation of n - l - 1; the current
    # superblock holds sys block start - 1 and env block n - start - 1
    win = ((start - 1, start + 1), (n - start - steps - 1, n - start - 1))
    full, part = d.fork(), d.fork(win)
    ra = full.sweep(0, start, start + steps, 1e-11, TRUN, 32, phase=2)
    rb = part.sweep(0, start, start + steps, 1e-11, TRUN, 32, phase=2)
    assert len(ra) == len(rb) == steps
    for x, y in zip(ra, rb):
        assert x.energy == y.energy and x.iters == y.iters and x.num_state == y.num_state, (x, y)
    with pytest.raises(HHError) as e:
        part.sweep(0, start + steps, start + steps + 1, 1e-11, TRUN, 32, phase=2)
    assert "window" in str(e.value)
    with pytest.raises(HHError):
        d.fork(((start + 2, start + 3), (1, n)))


def test_forced_fallback_is_flagged_and_correct(ctx):
    m = lat.Model(kind="tj", Lx=4, Ly=4, J=0.4, t=1.0, flux=0.125, n_elec=14)
    d, start = _prepared(ctx, m, 40)
    a, b = d.fork(), d.fork()
    a.set_solver(0)
    b.set_solver(2)
    b.set_precond(1)
    b.set_fallback(1e-6)   # no restart can cut the residual a millionfold
    ra = a.sweep(0, start, start + 2, 1e-11, TRUN, 40, phase=2)
    rb = b.sweep(0, start, start + 2, 1e-11, TRUN, 40, phase=2)
    assert all(r.fallback for r in rb if r.hpsi_fp64 > 2)
    assert any(r.fallback for r in rb)
    assert abs(ra[0].energy - rb[0].energy) <= 1e-9 * abs(ra[0].energy)


def test_trained_network_gain_over_the_method_without_it(ctx):
    """Train on triples from the first steps, then sweep the following steps
    from one state with the network, without it (fixed diagonal form) and
    with no preconditioner; the H.psi totals are printed (the gain) and the
    network may not cost more than 20 % over the form without it."""
    m = lat.Model(kind="tj", Lx=6, Ly=4, J=0.4, t=1.0, flux=0.125, n_elec=20)
    from hofstadter_hubbard_tsc_amd import Driver
    ctx.set_model(m)
    n = m.nsite
    kept = 120
    d = Driver(ctx, kept_max=kept, tol=1e-9, kept_max_warm=kept, seed=5)
    d.set_min_dim(0)
    d.warmup()
    bg = n // 2 - 1
    d.capture(16)
    d.set_solver(2)
    d.set_precond(1)
    d.sweep(0, bg, bg + 4, 1e-9, TRUN, kept, fresh=True, phase=1)
    losses, ns = d.train_precond(epochs=100, lr=1e-3)
    print("training: %d triples, loss %.4f -> %.4f" % (ns, losses[0], losses[-1]))
    assert losses[-1] < losses[0]
    tot = {}
    for name, mode in (("learned", 0), ("diagonal", 1), ("none", 2)):
        f = d.fork()
        f.set_solver(2)
        f.set_precond(mode)
        recs = f.sweep(0, bg + 4, n - 2, 1e-9, TRUN, kept, phase=2)
        tot[name] = sum(r.hpsi_fp64 + r.hpsi_fp32 for r in recs)
        assert all(np.isfinite(r.energy) for r in recs)
    print("H.psi over the steps: learned %d, diagonal %d, none %d (gain of the network over the fixed "
          "diagonal form: %.3f)" % (tot["learned"], tot["diagonal"], tot["none"], tot["diagonal"] / tot["learned"]))
    assert tot["learned"] <= 1.2 * tot["diagonal"]
