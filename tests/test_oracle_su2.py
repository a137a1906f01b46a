"""Pins for the oracle's 6j routine (oracle/su2.py)."""
import itertools
import math

import pytest
from sympy import Rational
from sympy.physics.wigner import wigner_6j

from oracle.su2 import w6j


def _args():
    # every argument pattern the DMRG recoupling uses: two of the six are a
    # tensor rank (1/2, 1) or a site spin (0, 1/2), the others block spins
    out = []
    for j1, j2 in itertools.product(range(0, 9), repeat=2):
        for k in (0, 1, 2):
            for s in (0, 1):
                for J in range(abs(j1 - s), j1 + s + 1, 2):
                    for Jp in range(abs(j2 - s), j2 + s + 1, 2):
                        out.append((j2, Jp, s, J, j1, k))
    return out


def test_w6j_matches_sympy():
    n = 0
    for a in _args():
        try:
            ref = float(wigner_6j(*[Rational(x, 2) for x in a]))
        except ValueError:      # sympy rejects non-integer triads: symbol is 0
            ref = 0.0
        assert abs(w6j(*a) - ref) < 1e-14, a
        n += 1
    assert n > 500


@pytest.mark.parametrize("a,b,c", [(1, 1, 0), (2, 1, 1), (4, 3, 1), (6, 6, 2), (7, 8, 1)])
def test_w6j_zero_argument_closed_form(a, b, c):
    # {a b c; b a 0} = (-1)^(a+b+c) / sqrt((2a+1)(2b+1))   (doubled args)
    if (a + b + c) % 2:
        pytest.skip("parity")
    val = w6j(a, b, c, b, a, 0)
    ref = (-1) ** ((a + b + c) // 2) / math.sqrt((a + 1) * (b + 1))
    assert abs(val - ref) < 1e-15


def test_w6j_orthogonality():
    # sum_x (2x+1)(2e+1) {a b x; c d e}{a b x; c d f} = delta_ef
    a, b, c, d = 3, 4, 5, 2
    for e in range(1, 9, 2):
        for f in range(1, 9, 2):
            s = sum((x + 1) * (e + 1) * w6j(a, b, x, c, d, e) * w6j(a, b, x, c, d, f)
                    for x in range(abs(a - b), a + b + 1, 2))
            ok_e = abs(a - d) <= e <= a + d and abs(b - c) <= e <= b + c
            expect = 1.0 if (e == f and ok_e) else 0.0
            assert abs(s - expect) < 1e-13, (e, f, s)


def test_w6j_triangle_violation_is_zero():
    assert w6j(1, 1, 6, 1, 1, 1) == 0.0
    assert w6j(2, 2, 2, 2, 2, 9) == 0.0
