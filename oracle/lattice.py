"""Lattice geometry, DMRG site maps and couplings (oracle, test infrastructure).

Restates the reference's lattice setup:
  * site label i = y + (x-1)*Ly, column by column (``Get_Lattice``,
    code/warm_up.f90:796-812; ``dmrg_setup`` loops, :347-666);
  * the superblock geometry of a step with sys block length ns and env block
    length ne (``dmrg_setup(ns, ns21)``, code/warm_up.f90:357-376): the sys
    block and its centre site occupy positions 1..ns+1, the env block and its
    centre site are mirrored from the right end, position totsite+1-i for env
    index i, with totsite = Ly * nxc and nxc = (nn-1)//Ly + 1 (at least 2);
  * square lattice: y bonds (with the periodic wrap only for Ly > 2, as the
    reference drops the duplicate rung of a 2-leg ladder, :413-419), x bonds;
    triangular lattice adds the (x,y)-(x+1,y-1) diagonal (:497-505).
Flux is the Landau gauge of the benchmark configurations: x bonds real, the
y bond from (x,y) to (x,y+1) carries t*exp(i 2 pi Phi x) for c^dag_(x,y)
c_(x,y+1) (x is the 1-based column); on the triangular lattice the diagonal
from (x+1,y-1) to (x,y) carries the same gauge's straight-line phase,
t*exp(-i 2 pi Phi (x + 1/2)) for c^dag_(x,y) c_(x+1,y-1), so every triangle
holds Phi/2.  The reference hard-codes its own phase pattern on the
triangular lattice instead (``get_phase``, code/quantum.f90:293-438), which
gauge="reference" restates (``_reference_gauge_bonds``); its
couplings are
jt = -sqrt(2) t, jd = -sqrt(3) J, jn = -J/4 (code/quantum.f90:93-122), which
are the reduced-tensor forms of the physical terms below.
"""
import cmath
import math
from dataclasses import dataclass, field


@dataclass
class Model:
    """Physical model: Hubbard (t, U) or t-J (t, J with -J/4 n n)."""
    kind: str = "hubbard"          # "hubbard" | "tj"
    Lx: int = 4
    Ly: int = 2
    lattice: str = "square"        # "square" | "triangular"
    periodic_y: bool = True
    t: float = 1.0
    U: float = 8.0
    J: float = 0.4
    flux: float = 0.0              # Phi per plaquette (Landau gauge)
    n_elec: int = 8
    gauge: str = "landau"          # "landau" | "reference" (get_phase, gauge = 1; triangular)
    theta: float = 0.0             # "reference": angle of the x bonds (the reference runs pi/2)

    @property
    def nsite(self):
        return self.Lx * self.Ly


def _xy(pos, Ly):
    return (pos - 1) // Ly + 1, (pos - 1) % Ly + 1


def lattice_bonds(model, ncol):
    """Unordered nearest-neighbour bonds of an ncol x Ly lattice (open in x).

    Returns a list of (u, v, kind, amp) with u < v; amp is the amplitude of
    c^dag_u c_v for hopping (its h.c. carries conj(amp)); kind is "y", "ywrap",
    "x" or "diag".
    """
    Ly = model.Ly
    out = []
    tw = 2.0 * math.pi * model.flux
    if getattr(model, "gauge", "landau") == "reference":
        return _reference_gauge_bonds(model, ncol)

    def hop_y(x):
        return -model.t * cmath.exp(1j * tw * x)

    def hop_d(x):
        # (x, y) <- (x+1, y-1): the straight-line Peierls phase of the same
        # Landau gauge, A_y = -2 pi flux x averaged over the bond (x + 1/2)
        return -model.t * cmath.exp(-1j * tw * (x + 0.5))

    for x in range(1, ncol + 1):
        for y in range(1, Ly + 1):
            u = y + (x - 1) * Ly
            # y bond within the column
            if y < Ly:
                out.append((u, u + 1, "y", hop_y(x)))
            elif Ly > 2 and model.periodic_y:
                # wrap bond (x,Ly)->(x,1): c^dag_(x,Ly) c_(x,1) has amp hop_y(x);
                # stored with u=(x,1) < v=(x,Ly): c^dag_u c_v = conj.
                v = 1 + (x - 1) * Ly
                out.append((v, u, "ywrap", hop_y(x).conjugate()))
            if x < ncol:
                out.append((u, u + Ly, "x", complex(-model.t)))
                if model.lattice == "triangular":
                    if y > 1:
                        out.append((u, u + Ly - 1, "diag", hop_d(x)))
                    elif Ly > 2 and model.periodic_y:
                        out.append((u, u + 2 * Ly - 1, "diag", hop_d(x)))
    return [(min(a, b), max(a, b), k, amp if a < b else amp.conjugate())
            for (a, b, k, amp) in out]


def _reference_gauge_bonds(model, ncol):
    """get_phase with gauge = 1 (code/quantum.f90:299-326) on the triangular
    lattice of dmrg_setup (code/warm_up.f90:469-505).  For site i = (x, y)
    (label y + (x-1) Ly) the reference sets pint(i, j) for its neighbours
    k = 1: j = (x, y+1) (y wraps), pint = phasey1 (-1)^x with phasey1 = -1;
    k = 3: j = (x+1, y), pint = exp(i theta);
    k = 5: j = (x+1, y-1) (y wraps), pint = phasey1 (-1)^(x+1);
    and pint(j, i) = conj(pint(i, j)).  The hopping amplitude of
    c^dag_i c_j is -t pint(i, j): the reference's term coefficient is
    jt pint(is, js) with jt = -sqrt(2) t the reduced form of -t
    (code/mat_vec.f90:3524, code/get_ham.f90:62)."""
    if model.lattice != "triangular":
        raise ValueError("the reference gauge is defined on the triangular lattice")
    Ly = model.Ly
    out = []
    for x in range(1, ncol + 1):
        for y in range(1, Ly + 1):
            i = y + (x - 1) * Ly
            p_y = -1.0 * (-1) ** x
            p_d = -1.0 * (-1) ** (x + 1)
            p_x = cmath.exp(1j * model.theta)
            if y < Ly:
                out.append((i, i + 1, "y", -model.t * p_y))
            elif Ly > 2 and model.periodic_y:
                out.append((i, 1 + (x - 1) * Ly, "ywrap", -model.t * p_y))
            if x < ncol:
                out.append((i, i + Ly, "x", -model.t * p_x))
                if y > 1:
                    out.append((i, i + Ly - 1, "diag", -model.t * p_d))
                elif Ly > 2 and model.periodic_y:
                    out.append((i, i + 2 * Ly - 1, "diag", -model.t * p_d))
    return [(min(a, b), max(a, b), k, complex(amp) if a < b else complex(amp).conjugate())
            for (a, b, k, amp) in out]


@dataclass
class Geometry:
    """Superblock geometry of one DMRG step (``dmrg_setup``)."""
    ns: int                        # sys block length
    ne: int                        # env block length
    ncol: int
    totsite: int
    sys_pos: list = field(default_factory=list)   # sys block site k (1..ns) -> position
    env_pos: list = field(default_factory=list)   # env block index k (1..ne) -> position
    sys_centre: int = 0
    env_centre: int = 0


def geometry(model, ns, ne):
    nn = ns + ne + 2
    ncol = (nn - 1) // model.Ly + 1
    if ncol == 1:
        ncol = 2
    tot = model.Ly * ncol
    g = Geometry(ns=ns, ne=ne, ncol=ncol, totsite=tot)
    g.sys_pos = [k for k in range(1, ns + 1)]
    g.sys_centre = ns + 1
    g.env_pos = [tot + 1 - k for k in range(1, ne + 1)]
    g.env_centre = tot - ne
    return g


def full_neighbours(model):
    """Neighbour sets of the full lattice (used for boundary-operator sets)."""
    nb = {i: set() for i in range(1, model.nsite + 1)}
    for u, v, _, _ in lattice_bonds(model, model.Lx):
        nb[u].add(v)
        nb[v].add(u)
    return nb


def boundary_sites(model, length):
    """Sites k (1..length, in the block's own index) that still have bonds to
    sites beyond the block in the full lattice (``dmrg_setup0``,
    code/warm_up.f90:311-341, keeps the same set plus neighbours of the last
    block site).  The env block uses the same table through the inversion
    i -> Nsite+1-i, which is a symmetry of the square and triangular lattices.
    """
    nb = full_neighbours(model)
    return [k for k in range(1, length + 1) if any(j > length for j in nb[k])]


def electrons_for_length(model, nsites):
    """Warm-up electron count (code/warm_up.f90:1386-1403): even floor of
    nsites/Nsite*N (+0.001), capped at N."""
    n = int(1.0 * nsites / model.nsite * model.n_elec + 0.001)
    n -= n % 2
    return min(n, model.n_elec)
