"""Sweep driver of the oracle (test infrastructure only).

Restates ``Get_Ground_State`` (code/quantum.f90:179-288) with trun_idx = 1:
infinite-system warm-up to Nsite/2 (``warm_up_point``,
code/warm_up.f90:1284-1491), a left-to-right half sweep from the centre and a
full right-to-left sweep (``left_to_right`` / ``right_to_left``,
code/warm_up.f90:1495-1849), then ``n_sweeps`` full sweeps.  Lanczos and
truncation thresholds follow the reference schedule: warm-up lanzero =
1000*tol and trun_err*100, first two half sweeps 15*tol and trun_err*15,
sweeps tol and trun_err, with trun_err = float32(1e-18) widened to double
(code/quantum.f90:9, :189-270).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import dmrg
from . import lattice as lat

TRUN_ERR = float(np.float32(1.0e-18))


@dataclass
class Schedule:
    kept_max: int = 64
    kept_max_warm: int = None          # kept_max during warm-up / first sweep
    n_sweeps: int = 2
    tol: float = 1e-10
    warm_tol_factor: float = 1000.0
    first_tol_factor: float = 15.0
    max_iter: int = 100
    seed: int = 1234

    def kept_min(self, kmax):
        return min(3 * kmax // 4, 3600)


@dataclass
class StepRecord:
    phase: str
    sys_len: int
    env_len: int
    energy: float
    iters: int
    num_state: int
    trunc_err: float
    entropy: float
    dim: int
    kept_per_sector: list = field(default_factory=list)


class Run:
    def __init__(self, model, sched):
        self.m = model
        self.s = sched
        self.sys_blocks = {}
        self.env_blocks = {}
        self.sys_trun = {}
        self.env_trun = {}
        self.records = []
        self.psi = None
        self.psi_ctx = None
        self.step = 0

    # -- one superblock solve ------------------------------------------------
    def _solve(self, sysb, envb, N, lanzero, psi0):
        sg = dmrg.StepGeometry(self.m, sysb.length, envb.length)
        se = dmrg.enlarge_in_geometry(sysb, self.m, "sys", sg)
        ee = dmrg.enlarge_in_geometry(envb, self.m, "env", sg)
        H = dmrg.Superblock(se, ee, sg, self.m, N)
        if psi0 is None:
            psi0 = dmrg.start_vector(H.sb.dim, self.s.seed + self.step)
        self.step += 1
        E, psi, it = dmrg.lanczos(H.matvec, psi0, lanzero, self.s.max_iter)
        return E, psi, it, se, ee, H

    def warm_up(self):
        m, s = self.m, self.s
        kmax = s.kept_max_warm or s.kept_max
        kmin = s.kept_min(kmax)
        lanzero = s.tol * s.warm_tol_factor
        terr = TRUN_ERR * 100
        sysb = dmrg.site_block(m, 0)
        envb = dmrg.site_block(m, 0)
        self.sys_blocks[1] = sysb
        self.env_blocks[1] = envb
        label = 1
        while label + 1 <= m.nsite // 2:
            N = lat.electrons_for_length(m, 2 * label + 2)
            E, psi, it, se, ee, H = self._solve(sysb, envb, N, lanzero, None)
            recs = []
            for side, en, blocks, truns in (("sys", se, self.sys_blocks, self.sys_trun),
                                            ("env", ee, self.env_blocks, self.env_trun)):
                rho = dmrg.density_matrix(H.sb, psi, side)
                if sum(en.basis.dims) <= kmax:
                    tr = dmrg.TruncResult(dmrg.identity_trun(en.basis), sum(en.basis.dims), 0.0, 0.0, [])
                else:
                    tr = dmrg.truncation(rho, kmin, kmax, terr)
                nb = dmrg.truncate_block(en, tr.trun, m, label + 1)
                blocks[label + 1] = nb
                truns[label + 1] = tr.trun
                recs.append(tr)
            sysb = self.sys_blocks[label + 1]
            envb = self.env_blocks[label + 1]
            self.records.append(StepRecord("warmup", label, label, E, it, recs[0].num_state,
                                           recs[0].err, recs[0].entropy, H.sb.dim, recs[0].kept_per_sector))
            label += 1

    def left_to_right(self, bg, end, lanzero, terr, kmax, fresh, phase):
        m = self.m
        kmin = self.s.kept_min(kmax)
        label = bg
        prev = None
        while label + 1 <= end:
            sysb = self.sys_blocks[label]
            envb = self.env_blocks[m.nsite - label - 2]
            if prev is None:
                psi0 = None if fresh else self.psi
            else:
                psi0 = None
            E, psi, it, se, ee, H = self._solve_with_guess(sysb, envb, m.n_elec, lanzero, psi0, prev, "LR")
            rho = dmrg.density_matrix(H.sb, psi, "sys")
            tr = dmrg.truncation(rho, kmin, kmax, terr)
            self.sys_blocks[label + 1] = dmrg.truncate_block(se, tr.trun, m, label + 1)
            self.sys_trun[label + 1] = tr.trun
            self.records.append(StepRecord(phase, label, envb.length, E, it, tr.num_state, tr.err,
                                           tr.entropy, H.sb.dim, tr.kept_per_sector))
            prev = (H, se, ee, psi, tr.trun)
            self.psi = psi
            label += 1

    def right_to_left(self, bg, end, lanzero, terr, kmax, phase):
        m = self.m
        kmin = self.s.kept_min(kmax)
        label = bg
        prev = None
        while label - 1 >= end:
            sysb = self.sys_blocks[label - 1]
            envb = self.env_blocks[m.nsite - label - 1]
            psi0 = self.psi if prev is None else None
            E, psi, it, se, ee, H = self._solve_with_guess(sysb, envb, m.n_elec, lanzero, psi0, prev, "RL")
            rho = dmrg.density_matrix(H.sb, psi, "env")
            tr = dmrg.truncation(rho, kmin, kmax, terr)
            self.env_blocks[m.nsite - label] = dmrg.truncate_block(ee, tr.trun, m, m.nsite - label)
            self.env_trun[m.nsite - label] = tr.trun
            self.records.append(StepRecord(phase, sysb.length, envb.length, E, it, tr.num_state, tr.err,
                                           tr.entropy, H.sb.dim, tr.kept_per_sector))
            prev = (H, se, ee, psi, tr.trun)
            self.psi = psi
            label -= 1

    def _solve_with_guess(self, sysb, envb, N, lanzero, psi0, prev, direction):
        """Build the superblock; the start vector is psi0 (carried over when
        the superblock is unchanged), the wave transformation of the previous
        step's ground state, or a fresh seeded vector."""
        sg = dmrg.StepGeometry(self.m, sysb.length, envb.length)
        se = dmrg.enlarge_in_geometry(sysb, self.m, "sys", sg)
        ee = dmrg.enlarge_in_geometry(envb, self.m, "env", sg)
        H = dmrg.Superblock(se, ee, sg, self.m, N)
        if prev is not None:
            pH, pse, pee, ppsi, ptrun = prev
            if direction == "LR":
                psi0 = dmrg.wave_trans_LR(pH.sb, pse, pee, ppsi, ptrun,
                                          self.env_trun[envb.length + 1], se, ee, H.sb)
            else:
                psi0 = dmrg.wave_trans_RL(pH.sb, pse, pee, ppsi, self.sys_trun[sysb.length + 1],
                                          ptrun, se, ee, H.sb)
        elif psi0 is not None and len(psi0) != H.sb.dim:
            psi0 = None
        if psi0 is None:
            psi0 = dmrg.start_vector(H.sb.dim, self.s.seed + self.step)
        self.step += 1
        E, psi, it = dmrg.lanczos(H.matvec, psi0, lanzero, self.s.max_iter)
        return E, psi, it, se, ee, H

    def run(self):
        m, s = self.m, self.s
        n = m.nsite
        kw = s.kept_max_warm or s.kept_max
        self.warm_up()
        self.left_to_right(n // 2 - 1, n - 2, s.tol * s.first_tol_factor, TRUN_ERR * 15, kw, True, "sweep0")
        self.right_to_left(n - 2, 1, s.tol * s.first_tol_factor, TRUN_ERR * 15, kw, "sweep0")
        for k in range(s.n_sweeps):
            self.left_to_right(1, n - 2, s.tol, TRUN_ERR, s.kept_max, False, "sweep%d" % (k + 1))
            self.right_to_left(n - 2, 1, s.tol, TRUN_ERR, s.kept_max, "sweep%d" % (k + 1))
        return self.records


def run(model, sched):
    return Run(model, sched).run()
