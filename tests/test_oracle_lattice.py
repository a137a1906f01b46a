"""The oracle's lattice: bond counts and the Landau-gauge flux through every
elementary plaquette (square) and triangle (triangular), read off the hopping
amplitudes as the phase an electron collects around the loop; the
reference's own gauge-1 phase pattern on the triangular lattice by its
triangle phases and, through exact diagonalisation, as a gauge of the Landau
field."""
import cmath
import math

import numpy as np
import pytest

from oracle import lattice as lat


def _amp(bonds, a, b):
    """amplitude of c^dag_a c_b (an electron hops b -> a)"""
    for (u, v, _, amp) in bonds:
        if (u, v) == (a, b):
            return amp
        if (u, v) == (b, a):
            return amp.conjugate()
    raise KeyError((a, b))


def _loop_phase(bonds, path):
    """Peierls phase collected along path[0] -> path[1] -> ... -> path[0]:
    the arg of the product of the amplitudes, each divided by -t"""
    z = 1.0
    for i in range(len(path)):
        src, dst = path[i], path[(i + 1) % len(path)]
        a = _amp(bonds, dst, src)
        z *= -a / abs(a)
    return cmath.phase(z)


def _site(x, y, Ly):
    return (y - 1) % Ly + 1 + (x - 1) * Ly


@pytest.mark.parametrize("lattice,Lx,Ly,periodic,count", [
    ("square", 4, 2, True, 4 * 1 + 3 * 2),          # ladder: no duplicate rung
    ("square", 4, 4, True, 4 * 4 + 3 * 4),          # cylinder
    ("square", 3, 4, False, 3 * 3 + 2 * 4),
    ("triangular", 4, 4, True, 4 * 4 + 3 * 4 + 3 * 4),
    ("triangular", 4, 3, False, 4 * 2 + 3 * 3 + 3 * 2),
])
def test_bond_counts(lattice, Lx, Ly, periodic, count):
    m = lat.Model(Lx=Lx, Ly=Ly, lattice=lattice, periodic_y=periodic, flux=0.1)
    assert len(lat.lattice_bonds(m, Lx)) == count


@pytest.mark.parametrize("flux", [0.125, 1.0 / 12, 0.3])
def test_square_plaquette_flux_is_uniform(flux):
    """Every plaquette (the periodic wrap included) encloses 2 pi flux."""
    m = lat.Model(Lx=5, Ly=4, flux=flux)
    bonds = lat.lattice_bonds(m, m.Lx)
    want = cmath.phase(cmath.exp(-2j * math.pi * flux))
    for x in range(1, m.Lx):
        for y in range(1, m.Ly + 1):
            loop = [_site(x, y, m.Ly), _site(x + 1, y, m.Ly), _site(x + 1, y + 1, m.Ly), _site(x, y + 1, m.Ly)]
            assert abs(_loop_phase(bonds, loop) - want) < 1e-12, (x, y)


@pytest.mark.parametrize("flux", [0.125, 1.0 / 16, 0.3])
def test_triangle_flux_is_half_the_plaquette(flux):
    """On the triangular lattice both triangles of every plaquette (periodic
    wrap included) enclose pi flux, counterclockwise as for the squares."""
    m = lat.Model(Lx=5, Ly=4, lattice="triangular", flux=flux)
    bonds = lat.lattice_bonds(m, m.Lx)
    want = cmath.phase(cmath.exp(-1j * math.pi * flux))
    for x in range(1, m.Lx):
        for y in range(1, m.Ly + 1):
            # the diagonal joins (x, y+1) and (x+1, y)
            lower = [_site(x, y, m.Ly), _site(x + 1, y, m.Ly), _site(x, y + 1, m.Ly)]
            upper = [_site(x + 1, y, m.Ly), _site(x + 1, y + 1, m.Ly), _site(x, y + 1, m.Ly)]
            assert abs(_loop_phase(bonds, lower) - want) < 1e-12, (x, y)
            assert abs(_loop_phase(bonds, upper) - want) < 1e-12, (x, y)


@pytest.mark.parametrize("theta", [math.pi / 2, 0.3, 1.1])
def test_reference_gauge_triangle_phases(theta):
    """get_phase's gauge-1 pattern (code/quantum.f90:299-326): counterclockwise,
    the upper triangle of every plaquette holds theta and the lower one
    pi - theta, periodic wrap included; at the reference's theta = pi/2 the
    field is uniform.  Every bond keeps |amp| = t, the bonds are those of the
    triangular lattice, and the loop around the even cylinder collects no
    phase."""
    m = lat.Model(Lx=5, Ly=4, lattice="triangular", gauge="reference", theta=theta)
    bonds = lat.lattice_bonds(m, m.Lx)
    landau = lat.lattice_bonds(lat.Model(Lx=5, Ly=4, lattice="triangular", flux=0.2), m.Lx)
    assert sorted((u, v) for u, v, _, _ in bonds) == sorted((u, v) for u, v, _, _ in landau)
    assert all(abs(abs(a) - m.t) < 1e-14 for _, _, _, a in bonds)
    for x in range(1, m.Lx):
        for y in range(1, m.Ly + 1):
            lower = [_site(x, y, m.Ly), _site(x + 1, y, m.Ly), _site(x, y + 1, m.Ly)]
            upper = [_site(x + 1, y, m.Ly), _site(x + 1, y + 1, m.Ly), _site(x, y + 1, m.Ly)]
            assert abs(cmath.exp(1j * _loop_phase(bonds, upper)) - cmath.exp(1j * theta)) < 1e-12, (x, y)
            assert abs(cmath.exp(1j * _loop_phase(bonds, lower)) - cmath.exp(1j * (math.pi - theta))) < 1e-12
    for x in range(1, m.Lx + 1):
        ring = [_site(x, y, m.Ly) for y in range(1, m.Ly + 1)]
        assert abs(cmath.exp(1j * _loop_phase(bonds, ring)) - 1.0) < 1e-12


def _lowest_singlets(m, k=3):
    from oracle import dmrg, ed
    sg = dmrg.StepGeometry(m, m.nsite // 2 - 1, m.nsite - m.nsite // 2 - 1)
    bonds = [(u, v) + sg.coupling(u, v) for (u, v) in sg.bonds]
    return ed.singlet_spectrum(m, list(range(1, m.nsite + 1)), bonds, m.n_elec, k=k)


def test_reference_gauge_is_a_gauge_of_the_landau_field():
    """At theta = pi/2 every triangle holds pi/2 and the cylinder loop
    nothing, as in the Landau gauge at Phi = -1/2 (the triangles hold
    -pi Phi): the two Hamiltonians are related by a gauge transformation, so
    their singlet spectra agree (exact diagonalisation of the triangular
    Hubbard cylinder 3 x 4); a different flux (Phi = 1/4) does not."""
    ref = lat.Model(kind="hubbard", Lx=3, Ly=4, lattice="triangular", gauge="reference", theta=math.pi / 2,
                    U=8.0, t=1.0, n_elec=6)
    lan = lat.Model(kind="hubbard", Lx=3, Ly=4, lattice="triangular", flux=-0.5, U=8.0, t=1.0, n_elec=6)
    other = lat.Model(kind="hubbard", Lx=3, Ly=4, lattice="triangular", flux=0.25, U=8.0, t=1.0, n_elec=6)
    a, b, c = _lowest_singlets(ref), _lowest_singlets(lan), _lowest_singlets(other)
    assert np.allclose(a, b, rtol=0, atol=1e-9), (a, b)
    assert np.abs(a - c).max() > 1e-3
