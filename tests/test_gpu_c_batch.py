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
