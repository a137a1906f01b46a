"""The product's matvec task compiler (hmerge + model_wave lowering, C++)
executed with numpy on the CPU must reproduce the oracle's H.psi on the same
blocks: every term type, both merge orientations, flux phases, Hubbard and
t-J."""
import numpy as np
import pytest

from hofstadter_hubbard_tsc_amd import _lib as L
from oracle import dmrg, driver
from oracle import lattice as lat

from . import hostpack, npexec


def _run_oracle(model, kept):
    run = driver.Run(model, driver.Schedule(kept_max=kept, n_sweeps=0, tol=1e-8))
    run.warm_up()
    n = model.nsite
    run.left_to_right(n // 2 - 1, n - 2, 1e-8, driver.TRUN_ERR, kept, True, "s0")
    run.right_to_left(n - 2, 1, 1e-8, driver.TRUN_ERR, kept, "s0")
    return run


def _check_step(model, run, ns):
    n = model.nsite
    sysb = run.sys_blocks[ns]
    envb = run.env_blocks[n - ns - 2]
    sg = dmrg.StepGeometry(model, sysb.length, envb.length)
    se = dmrg.enlarge_in_geometry(sysb, model, "sys", sg)
    ee = dmrg.enlarge_in_geometry(envb, model, "env", sg)
    H = dmrg.Superblock(se, ee, sg, model, model.n_elec)
    plan = hostpack.plan_matvec(model, sysb, envb, model.n_elec)
    try:
        lib = L.load()
        d = L.i64()
        npairs = L.i32()
        lib.hh_plan_super(plan, d, npairs, None, None, None, None, None)
        assert d.value == H.sb.dim
        nk = 1 if model.kind == "hubbard" else 3
        sys_arena = hostpack.pack(plan, 0, sysb, nk)
        env_arena = hostpack.pack(plan, 1, envb, nk)
        rng = np.random.default_rng(ns)
        psi = rng.normal(size=H.sb.dim) + 1j * rng.normal(size=H.sb.dim)
        ref = H.matvec(psi)
        t0, m0, s0 = npexec.fetch(plan, 0)
        t1, m1, s1 = npexec.fetch(plan, 1)
        t2, m2, _ = npexec.fetch(plan, 2)
        merged = np.zeros(max(s0, 1), np.complex128)
        X = np.zeros(max(s1, 1), np.complex128)
        out = np.zeros(H.sb.dim, np.complex128)
        bufs = [psi, out, sys_arena, env_arena, X, merged, None, None]
        npexec.run(t0, m0, bufs)
        npexec.run(t1, m1, bufs)
        npexec.run(t2, m2, bufs)
        err = np.abs(out - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < 1e-12, (ns, err)
        # the strided executor (the CPU baseline's) runs the same lists alike
        out2 = np.full(H.sb.dim, np.nan + 0j)
        bufs2 = [psi, out2, sys_arena, env_arena, np.zeros_like(X), np.zeros_like(merged), None, None]
        for tt, mm in ((t0, m0), (t1, m1), (t2, m2)):
            npexec.run_strided(tt, mm, bufs2)
        assert np.abs(out2 - out).max() <= 1e-13 * max(np.abs(out).max(), 1e-300), ns
    finally:
        L.load().hh_plan_destroy(plan)


@pytest.mark.parametrize("kind,Lx,Ly,flux,N,kept", [
    ("hubbard", 4, 2, 0.15, 6, 30),
    ("tj", 4, 2, 0.2, 6, 30),
    ("tj", 3, 4, 0.125, 10, 40),
    ("hubbard", 3, 3, 1.0 / 9, 8, 40),
])
def test_matvec_tasks_match_oracle(kind, Lx, Ly, flux, N, kept):
    model = lat.Model(kind=kind, Lx=Lx, Ly=Ly, U=8.0, J=0.4, t=1.0, flux=flux, n_elec=N)
    run = _run_oracle(model, kept)
    for ns in range(1, model.nsite - 2):
        _check_step(model, run, ns)


def test_matvec_tasks_match_oracle_triangular_flux():
    """The C++ task compiler's H.psi on a triangular cylinder with flux (the
    diagonal bonds' Landau phases) against the oracle's."""
    model = lat.Model(kind="tj", Lx=3, Ly=3, lattice="triangular", J=0.4, t=1.0, flux=1.0 / 6, n_elec=6)
    run = _run_oracle(model, 40)
    for ns in range(1, model.nsite - 2):
        _check_step(model, run, ns)


def test_matvec_tasks_match_oracle_reference_gauge():
    """The same on a triangular cylinder in get_phase's own gauge-1 pattern
    (theta = pi/2, code/quantum.f90:299-326): every bond's amplitude from the
    product's lattice generator enters the tasks as the oracle's does."""
    import math
    model = lat.Model(kind="tj", Lx=3, Ly=4, lattice="triangular", J=0.4, t=1.0, gauge="reference",
                      theta=math.pi / 2, n_elec=10)
    run = _run_oracle(model, 40)
    for ns in range(1, model.nsite - 2):
        _check_step(model, run, ns)
