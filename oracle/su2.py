"""SU(2) recoupling coefficients for the oracle (test infrastructure only).

Follows the Racah formula used by the reference's ``W6JS``
(code/update.f90:1665-1716): all arguments are doubled integers (2j), the
triangle conditions and parity checks are those of the reference, and the
result is the Wigner 6j symbol {j1 j2 j3; l1 l2 l3}.  Here the sum and the
triangle coefficients are evaluated exactly with Python integers / Fractions
and rounded to double once at the end.
"""
from fractions import Fraction
from functools import lru_cache
from math import factorial, sqrt


def _tri_ok(a, b, c):
    """Doubled-integer triangle rule |a-b| <= c <= a+b with a+b+c even."""
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def _delta(a, b, c):
    """Triangle coefficient Delta(abc)^2 as an exact Fraction (doubled args)."""
    return Fraction(factorial((a + b - c) // 2) * factorial((a - b + c) // 2)
                    * factorial((-a + b + c) // 2), factorial((a + b + c) // 2 + 1))


@lru_cache(maxsize=None)
def w6j(j1, j2, j3, l1, l2, l3):
    """Wigner 6j {j1 j2 j3; l1 l2 l3}, doubled-integer arguments.

    Reference: ``W6JS`` (code/update.f90:1665).  Zero when any of the four
    triads (j1 j2 j3), (j1 l2 l3), (l1 j2 l3), (l1 l2 j3) violates the
    triangle rule.
    """
    if not (_tri_ok(j1, j2, j3) and _tri_ok(j1, l2, l3)
            and _tri_ok(l1, j2, l3) and _tri_ok(l1, l2, j3)):
        return 0.0
    s1 = (j1 + j2 + j3) // 2
    s2 = (j1 + l2 + l3) // 2
    s3 = (l1 + j2 + l3) // 2
    s4 = (l1 + l2 + j3) // 2
    p1 = (j1 + j2 + l1 + l2) // 2
    p2 = (j2 + j3 + l2 + l3) // 2
    p3 = (j3 + j1 + l3 + l1) // 2
    total = 0
    for t in range(max(s1, s2, s3, s4), min(p1, p2, p3) + 1):
        den = (factorial(t - s1) * factorial(t - s2) * factorial(t - s3)
               * factorial(t - s4) * factorial(p1 - t) * factorial(p2 - t)
               * factorial(p3 - t))
        term = Fraction(factorial(t + 1), den)
        total += -term if t % 2 else term
    theta = _delta(j1, j2, j3) * _delta(j1, l2, l3) * _delta(l1, j2, l3) * _delta(l1, l2, j3)
    # total * sqrt(theta) = sign(total) * sqrt(total^2 * theta); the exact
    # Fraction is rounded to double once (float(Fraction) rounds correctly).
    val = sqrt(float(total * total * theta))
    return val if total >= 0 else -val


def phase(two_x):
    """(-1)^x for a doubled integer 2x that is even (x integer)."""
    assert two_x % 2 == 0, two_x
    return -1.0 if (two_x // 2) % 2 else 1.0
