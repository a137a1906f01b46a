"""The oracle's lattice: bond counts and the Landau-gauge flux through every
elementary plaquette (square) and triangle (triangular), read off the hopping
amplitudes as the phase an electron collects around the loop."""
import cmath
import math

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
