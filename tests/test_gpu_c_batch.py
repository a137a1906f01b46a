"""GPU parity of the batched H.psi (several superblocks, one grouped launch
per stage): every member against its own single-superblock H.psi and, for
two members, against the oracle matvec on the same (synthetic) blocks; and
of the batched Lanczos against each member's own Lanczos, with and without
the HIP-graph replay of its grouped H.psi."""
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
        ctx.batch([sb] * 33)
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


def _singles(ctx, members, starts, tol, max_iter):
    out = []
    for (sysb, envb, sb, _), p0 in zip(members, starts):
        p = p0.clone()
        e, it = ctx.lanczos(sb, p, tol, max_iter)
        conv, _, _ = ctx.lanczos_status()
        out.append((e, it, conv, p))
    return out


def test_batch_lanczos_matches_each_member_alone(ctx):
    """hh_batch_lanczos of five members of different sizes (so that they stop
    at different iterations) against hh_lanczos of each member alone from the
    same start vector: energy 1e-10 relative, ground-state overlap within 1e-8
    of 1, iteration counts within 2.  The HIP-graph replay of the grouped
    H.psi and the kernel-by-kernel launches give the same result bit for bit
    (the graph is replayed once per iteration with the solver's own kernels
    and host reads in between).  A batch of one member runs the same H.psi
    task list as the member alone: its energy and iteration count equal
    hh_lanczos's."""
    import torch
    from hofstadter_hubbard_tsc_amd import synth
    members = [synth.make_member(ctx, b, m_lo=200, m_hi=420) for b in range(5)]
    starts = [m[3].clone() for m in members]
    tol, cap = 1e-9, 80
    ref = _singles(ctx, members, starts, tol, cap)
    batch = ctx.batch([m[2] for m in members])
    res = {}
    for graph in (True, False):
        batch.set_graph(graph)
        psis = [p.clone() for p in starts]
        e0, it, cv = ctx.batch_lanczos(batch, psis, tol, cap)
        res[graph] = (e0, it, cv, psis)
    eg, ig, cg, pg = res[True]
    en, in_, cn, pn = res[False]
    assert np.array_equal(eg, en) and np.array_equal(ig, in_) and np.array_equal(cg, cn)
    for a, b in zip(pg, pn):
        assert torch.equal(a, b)
    print("iterations batched", ig.tolist(), "alone", [r[1] for r in ref])
    assert len(set(ig.tolist())) > 1, "the members should stop at different iterations"
    for i, (e, it, conv, p) in enumerate(ref):
        assert abs(eg[i] - e) <= 1e-10 * abs(e), (i, eg[i], e)
        assert abs(int(ig[i]) - it) <= 2, (i, ig[i], it)
        ov = abs(torch.vdot(pg[i], p).item())
        assert ov >= 1 - 1e-8, (i, ov)
        assert abs(torch.linalg.vector_norm(pg[i]).item() - 1.0) < 1e-12
    one = ctx.batch([members[2][2]])
    p1 = [starts[2].clone()]
    e1, i1, _ = ctx.batch_lanczos(one, p1, tol, cap)
    print("one-member batch: E %.17g alone %.17g, iterations %d / %d" % (e1[0], ref[2][0], i1[0], ref[2][1]))
    assert abs(e1[0] - ref[2][0]) <= 1e-12 * abs(ref[2][0]) and abs(int(i1[0]) - ref[2][1]) <= 1


def test_batch_holds_32_members(ctx):
    """32 members in one grouped H.psi and one batched Lanczos (the slot
    index of member 31 is the last one a uint8 holds); 33 are refused."""
    import torch
    from hofstadter_hubbard_tsc_amd import HHError, synth
    members = [synth.make_member(ctx, b, m_lo=40, m_hi=80) for b in range(32)]
    with pytest.raises(HHError):
        ctx.batch([members[0][2]] * 33)
    batch = ctx.batch([m[2] for m in members])
    outs = [torch.zeros_like(m[3]) for m in members]
    ctx.batch_model_wave(batch, [m[3] for m in members], outs)
    for (sysb, envb, sb, psi), yb in zip(members, outs):
        y1 = torch.zeros_like(psi)
        ctx.model_wave(sb, psi, y1)
        assert (yb - y1).abs().max().item() <= 1e-13 * y1.abs().max().item()
    psis = [m[3].clone() for m in members]
    e0, it, cv = ctx.batch_lanczos(batch, psis, 1e-9, 60)
    for i in (0, 31):
        p = members[i][3].clone()
        e, _ = ctx.lanczos(members[i][2], p, 1e-9, 60)
        assert abs(e0[i] - e) <= 1e-10 * abs(e), (i, e0[i], e)


def test_autotuned_launches_give_the_same_hpsi(ctx):
    """With autotuning on, binding a superblock times its two H.psi stages
    under the candidate launch cuts and keeps the fastest; the tuned H.psi
    equals the fixed cut's (1e-13, only the order of split partial sums can
    differ); binding a superblock of the same shape statistics again is a
    cache hit without timing; a tuned batch's grouped H.psi equals each
    member's own."""
    import torch
    from hofstadter_hubbard_tsc_amd import synth
    sysb, envb, sb0, psi = synth.make_member(ctx, 3, m_lo=300, m_hi=420)
    y0 = torch.zeros_like(psi)
    ctx.model_wave(sb0, psi, y0)
    before = ctx.autotune_stats()
    ctx.set_autotune(True)
    try:
        sb1 = ctx.bind(sysb, envb, synth.N_TARGET)
        st = ctx.autotune_stats()
        assert st["tunes"] - before["tunes"] + st["hits"] - before["hits"] == 2
        y1 = torch.zeros_like(psi)
        ctx.model_wave(sb1, psi, y1)
        assert (y1 - y0).abs().max().item() <= 1e-13 * y0.abs().max().item()
        sb2 = ctx.bind(sysb, envb, synth.N_TARGET)
        st2 = ctx.autotune_stats()
        assert st2["tunes"] == st["tunes"] and st2["hits"] == st["hits"] + 2
        members = [synth.make_member(ctx, b, m_lo=200, m_hi=300) for b in range(4)]
        batch = ctx.batch([m[2] for m in members])
        outs = [torch.zeros_like(m[3]) for m in members]
        ctx.batch_model_wave(batch, [m[3] for m in members], outs)
        for (s_, e_, sb, p), yb in zip(members, outs):
            y = torch.zeros_like(p)
            ctx.model_wave(sb, p, y)
            assert (yb - y).abs().max().item() <= 1e-13 * y.abs().max().item()
        print("autotune:", ctx.autotune_stats())
    finally:
        ctx.set_autotune(False)
