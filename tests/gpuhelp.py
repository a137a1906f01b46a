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


def oracle_block(ctx, blk, model):
    """Product block (device arena) -> oracle Block with the same data."""
    from oracle import dmrg as odmrg
    arena = blk.download()
    N, S2, d, h = blk.sectors()
    sites = blk.sites()
    nk = 1 if model.kind == "hubbard" else 3
    H = odmrg.Op(0, 0, False)
    for i, dim in enumerate(d):
        H.blk[(i, i)] = arena[h[i]:h[i] + dim * dim].reshape((dim, dim), order="F").copy()
    ops = {}
    for si, k in enumerate(sites):
        ops[int(k)] = {}
        for kind in range(nk):
            base = ctx.op_offset(blk, si, kind)
            bra, ket, off = ctx.layout(blk, kind)
            op = odmrg.Op([1, 2, 0][kind], [1, 0, 0][kind], kind == 0)
            for b_, k_, o in zip(bra, ket, off):
                r, c = d[b_], d[k_]
                op.blk[(int(b_), int(k_))] = arena[base + o: base + o + r * c].reshape((r, c), order="F").copy()
            ops[int(k)][KIND_NAMES[kind]] = op
    _, _, _, length = blk.info()
    return odmrg.Block(length, list(zip(N.tolist(), S2.tolist())), d.tolist(), H, ops)


def oracle_superblock(ctx, sysb, envb, model, n_elec):
    """Oracle H of the superblock bound from two product blocks."""
    from oracle import dmrg as odmrg
    so, eo = oracle_block(ctx, sysb, model), oracle_block(ctx, envb, model)
    sg = odmrg.StepGeometry(model, so.length, eo.length)
    return odmrg.Superblock(odmrg.enlarge_in_geometry(so, model, "sys", sg),
                            odmrg.enlarge_in_geometry(eo, model, "env", sg), sg, model, n_elec)


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
