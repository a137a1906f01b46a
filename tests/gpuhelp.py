"""Helpers for the GPU tests: move oracle blocks into the product and back."""
import numpy as np

KIND_NAMES = ["cdag", "S", "n"]


def upload_block(ctx, block, side, nkinds):
    sectors = [(q[0], q[1], d) for q, d in zip(block.sectors, block.dims)]
    sites = sorted(block.ops.keys())
    b = ctx.create_block(side, block.length, sectors, sites)
    size, nsec, nsites, _ = b.info()
    N, S2, d, h = b.sectors()
    arena = np.zeros(max(size, 1), np.complex128)
    for i, dim in enumerate(d):
        arena[h[i]:h[i] + dim * dim] = block.H.blk[(i, i)].reshape(-1, order="F")
    for si, k in enumerate(sites):
        for kind in range(nkinds):
            base = ctx.op_offset(b, si, kind)
            bra, ket, off = ctx.layout(b, kind)
            op = block.ops[k][KIND_NAMES[kind]]
            for b_, k_, o in zip(bra, ket, off):
                m = op.blk.get((int(b_), int(k_)))
                if m is not None:
                    arena[base + o: base + o + m.size] = m.reshape(-1, order="F")
    b.upload(arena[:size] if size else arena)
    return b


def block_h_spectra(ctx, b):
    """Per-sector eigenvalues of H (gauge invariant)."""
    arena = b.download()
    N, S2, d, h = b.sectors()
    out = {}
    for i, dim in enumerate(d):
        m = arena[h[i]:h[i] + dim * dim].reshape((dim, dim), order="F")
        out[(int(N[i]), int(S2[i]))] = np.sort(np.linalg.eigvalsh(0.5 * (m + m.conj().T)))
    return out


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()
