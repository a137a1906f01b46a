"""GPU parity of the batched H.psi (several superblocks, one grouped launch
per stage): every member against its own single-superblock H.psi and, for
two members, against the oracle matvec on the same (synthetic) blocks."""
import numpy as np
import pytest

from oracle import lattice as lat

from .gpuhelp import oracle_superblock

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a device"
    from hofstadter_hubbard_tsc_amd import Context
    c = Context(0)
    yield c
    c.close()


def test_batch_matches_single_and_oracle(ctx):
    import torch
    from hofstadter_hubbard_tsc_amd import synth
    members = [synth.make_member(ctx, b, m_lo=240, m_hi=420) for b in range(5)]
    batch = ctx.batch([m[2] for m in members])
    info = batch.info()
    assert info["dim"] == sum(m[2].dim for m in members)
    outs = [torch.zeros_like(m[3]) for m in members]
    ctx.batch_model_wave(batch, [m[3] for m in members], outs)
    for (sysb, envb, sb, psi), yb in zip(members, outs):
        y1 = torch.zeros_like(psi)
        ctx.model_wave(sb, psi, y1)
        err = (yb - y1).abs().max().item() / y1.abs().max().item()
        assert err < 1e-13, err
    # a second call reuses the batch with new vectors
    psi2 = [torch.flip(m[3], [0]).contiguous() for m in members]
    outs2 = [torch.zeros_like(p) for p in psi2]
    ctx.batch_model_wave(batch, psi2, outs2)
    for (sysb, envb, sb, psi), p, yb in zip(members, psi2, outs2):
        y1 = torch.zeros_like(p)
        ctx.model_wave(sb, p, y1)
        assert (yb - y1).abs().max().item() <= 1e-13 * y1.abs().max().item()
    # oracle parity for the smallest and the largest member
    dims = [m[2].dim for m in members]
    for b in (int(np.argmin(dims)), int(np.argmax(dims))):
        sysb, envb, sb, psi = members[b]
        mp = synth.model_params((b + 1) / 64.0)
        model = lat.Model(kind="hubbard", Lx=mp.Lx, Ly=mp.Ly, U=mp.U, t=mp.t, flux=mp.flux, n_elec=mp.n_elec)
        H = oracle_superblock(ctx, sysb, envb, model, mp.n_elec)
        assert H.sb.dim == sb.dim
        ref = H.matvec(psi.cpu().numpy())
        got = outs[b].cpu().numpy()
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12, b


def test_batch_rejects_bad_arguments(ctx):
    from hofstadter_hubbard_tsc_amd import HHError, synth
    sysb, envb, sb, psi = synth.make_member(ctx, 0, m_lo=240, m_hi=260)
    with pytest.raises(HHError):
        ctx.batch([sb] * 32)
    batch = ctx.batch([sb])
    with pytest.raises(HHError):
        ctx.batch_model_wave(batch, [psi], [psi])


def test_full_size_batched_workload(ctx):
    """The batched H.psi workload at full size: 16 members with m drawn in
    [500, 4000] (the benchmark's own members, blocks up to 1024-dimensional
    sectors).  Every member of one grouped call equals its own
    single-superblock H.psi; the grouped call is linear (the partial sums of
    one call leave nothing behind for the next); the oracle matvec on the same
    blocks agrees for the smallest and the largest member."""
    import time

    import torch
    from hofstadter_hubbard_tsc_amd import synth
    members = synth.make_batch(ctx)
    dims = [m[2].dim for m in members]
    print(f"members {len(members)}, total dim {sum(dims)}, largest {max(dims)}", flush=True)
    batch = ctx.batch([m[2] for m in members])
    outs = [torch.zeros_like(m[3]) for m in members]
    ctx.batch_model_wave(batch, [m[3] for m in members], outs)
    for b, ((sysb, envb, sb, psi), yb) in enumerate(zip(members, outs)):
        y1 = torch.zeros_like(psi)
        ctx.model_wave(sb, psi, y1)
        err = (yb - y1).abs().max().item() / y1.abs().max().item()
        assert err < 1e-13, (b, err)
    # linearity: H(a psi + c phi) = a H psi + c H phi, phi the reversed psi
    a, c = complex(0.6, -0.3), complex(-0.2, 0.9)
    phi = [torch.flip(m[3], [0]).contiguous() for m in members]
    hphi = [torch.zeros_like(p) for p in phi]
    ctx.batch_model_wave(batch, phi, hphi)
    mix = [a * m[3] + c * p for m, p in zip(members, phi)]
    hmix = [torch.zeros_like(x) for x in mix]
    ctx.batch_model_wave(batch, mix, hmix)
    for b, (hm, y, hp) in enumerate(zip(hmix, outs, hphi)):
        ref = a * y + c * hp
        err = (hm - ref).abs().max().item() / ref.abs().max().item()
        assert err < 1e-13, (b, err)
    for b in (int(np.argmin(dims)), int(np.argmax(dims))):
        t0 = time.time()
        sysb, envb, sb, psi = members[b]
        mp = synth.model_params((b + 1) / 64.0)
        model = lat.Model(kind="hubbard", Lx=mp.Lx, Ly=mp.Ly, U=mp.U, t=mp.t, flux=mp.flux, n_elec=mp.n_elec)
        H = oracle_superblock(ctx, sysb, envb, model, mp.n_elec)
        assert H.sb.dim == sb.dim
        ref = H.matvec(psi.cpu().numpy())
        got = outs[b].cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"member {b}: dim {sb.dim}, oracle {time.time() - t0:.1f} s, rel err {err:.2e}", flush=True)
        assert err < 1e-12, (b, err)
