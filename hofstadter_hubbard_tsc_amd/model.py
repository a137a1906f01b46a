"""Model parameters and the named workloads of this library.

`Model` carries what `hh_set_model` takes (include/hhdmrg.h
`hh_model_params`): the couplings and lattice of program Quantum
(code/quantum.f90:93-160) and dmrg_setup (code/warm_up.f90:347-666), with
the flux in the Landau gauge (x bonds real, y bonds t exp(i 2 pi flux x)) or,
on the triangular lattice, in get_phase's own gauge-1 pattern
(code/quantum.f90:299-326) with x-bond angle theta.

`WORKLOADS` names the configurations the benchmark and the parity tests run,
each with its sweep schedule (Get_Ground_State, code/quantum.f90:179-288):

* c1: Hubbard 16 x 2 ladder (open in x and y), U = 8, 28 electrons, m = 256,
  warm-up + two warm-up half sweeps + 6 sweeps, Lanczos tolerance 1e-10;
* c2: flux t-J Ly = 4 cylinder x Lx = 32, J = 0.4, 112 electrons (hole
  doping 1/8), flux 1/8 per plaquette, m ramped 500 -> 2000 over 10 sweeps;
* c4: Hubbard Ly = 6 cylinder x Lx = 24, U = 8, 126 electrons (density
  7/8), flux 1/12, m = 4000;
* c5: t-J Ly = 6 cylinder x Lx = 48, J = 0.33, 264 electrons (doping 1/12),
  flux 1/16, m = 8000.
"""
from dataclasses import dataclass


@dataclass
class Model:
    kind: str = "hubbard"        # "hubbard" | "tj"
    Lx: int = 4
    Ly: int = 2
    lattice: str = "square"      # "square" | "triangular"
    periodic_y: bool = True      # the y wrap exists only for Ly > 2
    t: float = 1.0
    U: float = 8.0
    J: float = 0.4
    flux: float = 0.0            # flux quanta per plaquette (Landau gauge)
    n_elec: int = 8
    gauge: str = "landau"        # "landau" | "reference" (get_phase, gauge 1; triangular)
    theta: float = 0.0           # "reference": phase of the x bonds (the reference runs pi/2)

    @property
    def nsite(self):
        return self.Lx * self.Ly


@dataclass
class Workload:
    name: str
    model: Model
    kept_max: int                # m of the finite sweeps (top of the ramp)
    kept_max_warm: int           # m of the warm-up
    n_sweeps: int
    tol: float                   # Lanczos tolerance (lanzero) of the finite sweeps
    description: str


WORKLOADS = {
    "c1": Workload("c1", Model(kind="hubbard", Lx=16, Ly=2, periodic_y=False, t=1.0, U=8.0, flux=0.0, n_elec=28),
                   256, 256, 6, 1e-10,
                   "Hubbard 16x2 ladder (open x and y), t=1, U=8, N=28, S=0, no flux, m=256 SU(2) multiplets, "
                   "warm-up + 2 warm-up half sweeps + 6 sweeps, Lanczos tol 1e-10"),
    "c2": Workload("c2", Model(kind="tj", Lx=32, Ly=4, periodic_y=True, t=1.0, J=0.4, flux=1.0 / 8, n_elec=112),
                   2000, 500, 10, 1e-8,
                   "flux t-J cylinder Ly=4 x Lx=32, N=112, S=0, flux 1/8 (Landau), t=1, J=0.4, "
                   "m ramped 500 -> 2000 over 10 sweeps"),
    "c4": Workload("c4", Model(kind="hubbard", Lx=24, Ly=6, periodic_y=True, t=1.0, U=8.0, flux=1.0 / 12,
                               n_elec=126),
                   4000, 4000, 3, 1e-8,
                   "Hubbard cylinder Ly=6 x Lx=24, t=1, U=8, N=126 (density 7/8), S=0, flux 1/12 (Landau), "
                   "m=4000 SU(2) multiplets"),
    "c5": Workload("c5", Model(kind="tj", Lx=48, Ly=6, periodic_y=True, t=1.0, J=0.33, flux=1.0 / 16, n_elec=264),
                   8000, 8000, 1, 1e-8,
                   "t-J cylinder Ly=6 x Lx=48, t=1, J=0.33, N=264 (doping 1/12), S=0, flux 1/16 (Landau), "
                   "m=8000 SU(2) multiplets"),
}
