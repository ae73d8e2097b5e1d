"""The device sweeps divide by tabulated pivots as q0 = n r, e = fma(-b, q0,
n), q = fma(e, r, q0) with r = 1/b correctly rounded (Markstein's final
step).  This checks on the CPU, with exact rational arithmetic standing in
for the fused multiply-add, that the result is the same double as n / b for
the magnitudes the pressure and Crank-Nicolson systems produce."""
from fractions import Fraction

import numpy as np


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # one rounding


def _quotient(n, b):
    r = 1.0 / b
    q0 = n * r
    e = _fma(-b, q0, n)
    return _fma(e, r, q0)


def test_recip_quotient_equals_division():
    rng = np.random.default_rng(11)
    bad = 0
    for _ in range(20000):
        n = float(rng.standard_normal() * 10.0 ** rng.integers(-12, 12))
        b = float(rng.standard_normal() * 10.0 ** rng.integers(-8, 8))
        if b == 0.0:
            continue
        bad += _quotient(n, b) != n / b
    assert bad == 0


def test_recip_quotient_hard_cases():
    # quotients close to a rounding boundary and exact quotients
    cases = [(1.0, 3.0), (2.0, 3.0), (10.0, 7.0), (1.0, 49.0), (5.0, 5.0), (-7.5, 2.5),
             (1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53), (3.0, 1.0 + 2.0 ** -52)]
    for n, b in cases:
        assert _quotient(n, b) == n / b, (n, b)
