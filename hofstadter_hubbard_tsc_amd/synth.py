"""Synthetic superblocks for the batched H.psi benchmark.

The workload: independent Hubbard superblocks (site multiplets (N, 2S) =
(0,0), (1,1), (2,0)) on a Ly = 4 cylinder of 16 x 4 sites, 64 electrons,
total S = 0, each at its own flux.  Member b draws everything from a Philox
counter-based stream keyed by ``seed_base + b``:

* one total multiplet count m_b, uniform integer in [m_lo, m_hi], for both
  sides;
* 120 sectors per side, labels (N, 2S) on N = Nc-7 .. Nc+7 with 2S = N mod 2,
  N mod 2 + 2, ..., N mod 2 + 14, around Nc = 31 (the environment mirrors
  the system about the 64-electron target);
* raw dimensions from LogNormal(ln 48, 0.8), scaled to sum to m_b, rounded,
  clipped to [1, 1024], the residual put on the largest sector;
* every arena element (block H and boundary-site operators) complex Gaussian
  N(0,1) + i N(0,1), scaled by 0.1, drawn on the device from a generator
  seeded with the same key; each sector's block H is then replaced by its
  Hermitian part (H + H^dag) / 2, so that the superblock H is Hermitian (the
  hopping terms carry their own h.c.) and has a ground state;
* psi complex Gaussian, unit norm.

The 6j recoupling coefficients and flux phases come from the library's own
task compiler at bind time, so the sizes and the term structure are those
of a real step of the same lattice.
"""
from types import SimpleNamespace

import numpy as np

LX, LY, N_TARGET, BLOCK_LEN = 16, 4, 64, 31


def sector_table(rng, m_total, nc):
    """120 (N, 2S, dim) sectors around N = nc with dims summing to m_total."""
    labels = [(n, n % 2 + 2 * k) for n in range(nc - 7, nc + 8) for k in range(8)]
    raw = rng.lognormal(np.log(48.0), 0.8, len(labels))
    d = np.clip(np.rint(raw / raw.sum() * m_total), 1, 1024).astype(np.int64)
    d[int(np.argmax(d))] += m_total - int(d.sum())
    if d.min() < 1:   # the residual can only be negative by a few units
        d = np.maximum(d, 1)
    return [(n, s2, int(x)) for (n, s2), x in zip(labels, d)]


def model_params(flux):
    """Model description accepted by Context.set_model."""
    return SimpleNamespace(kind="hubbard", Lx=LX, Ly=LY, lattice="square", periodic_y=True, t=1.0, U=8.0, J=0.0,
                           flux=flux, n_elec=N_TARGET)


def make_member(ctx, b, seed_base=0x5EED0000, m_lo=500, m_hi=4000, device="cuda"):
    """(sys block, env block, superblock, psi) of member b."""
    import torch
    key = seed_base + b
    rng = np.random.Generator(np.random.Philox(key=key))
    ctx.set_model(model_params(flux=(b + 1) / 64.0))
    m = int(rng.integers(m_lo, m_hi + 1))
    sys_secs = sector_table(rng, m, BLOCK_LEN)
    env_secs = sector_table(rng, m, N_TARGET - 2 - BLOCK_LEN)
    sites = ctx.boundary_sites(BLOCK_LEN)
    gen = torch.Generator(device=device)
    gen.manual_seed(key)
    blocks = []
    for side, secs in ((0, sys_secs), (1, env_secs)):
        blk = ctx.create_block(side, BLOCK_LEN, secs, sites)
        size = blk.info()[0]
        a = torch.randn(size, dtype=torch.complex128, device=device, generator=gen) * 0.1
        _, _, dims, hoff = blk.sectors()
        for d, o in zip(dims.tolist(), hoff.tolist()):
            h = a[o:o + d * d].view(d, d)
            h.copy_(0.5 * (h + h.conj().transpose(0, 1)))
        blk.upload(a)
        del a
        blocks.append(blk)
    sb = ctx.bind(blocks[0], blocks[1], N_TARGET)
    psi = torch.randn(sb.dim, dtype=torch.complex128, device=device, generator=gen)
    psi /= torch.linalg.vector_norm(psi)
    return blocks[0], blocks[1], sb, psi


def make_batch(ctx, nsb=16, **kw):
    """nsb members (see make_member); returns their tuples."""
    return [make_member(ctx, b, **kw) for b in range(nsb)]
