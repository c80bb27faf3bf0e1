"""The marker kernels divide by grid steps as q = RN(a*rb) followed by two
remainder corrections with rb = RN(1/b) (rn_div in csrc/ebg_advect.hip).
This test evaluates that exact sequence with exact rational arithmetic (each
fma rounded once, as the hardware does) and checks it against the correctly
rounded quotient on random, near-tie and extreme-significand operands, within
the numerator range the kernel admits ([2^-900, 2^900] and zero) and the
positive divisors it admits ([2^-100, 2^100]); the sign of a zero numerator is
restored by copying the numerator's sign onto the result."""
import math
from fractions import Fraction

import numpy as np


def _fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def rn_div(a: float, b: float, rb: float) -> float:
    q = a * rb
    r = _fma(-b, q, a)
    q = _fma(r, rb, q)
    r = _fma(-b, q, a)
    q = _fma(r, rb, q)
    return math.copysign(q, a)


def _cases(rng):
    bs = [1500.0, 1200.0, 6.0, 3.0, 7.0, 0.1, 1.0 / 256.0, 1500.000000001,
          2.0 - 2.0 ** -52, 2.0 ** 20 * (2.0 - 2.0 ** -52), 1.0 + 2.0 ** -52]
    bs += list(rng.uniform(1e-3, 1e6, 30))
    bs += list(2.0 ** rng.uniform(-100, 100, 20))
    for b in bs:
        rb = 1.0 / b
        avals = list(rng.uniform(-2e6, 2e6, 150))
        avals += list(rng.uniform(-1.0, 1.0, 50) * b)
        for k in rng.integers(-50, 50, 40):  # a near k*b: quotient near an integer
            a0 = float(k) * b
            avals += [a0, np.nextafter(a0, np.inf), np.nextafter(a0, -np.inf)]
        for k in rng.integers(1, 2 ** 40, 20):  # quotients near a half-ulp
            q = float(k) + 0.5
            a0 = q * b
            avals += [a0, np.nextafter(a0, np.inf), np.nextafter(a0, -np.inf)]
        avals += [0.0, -0.0, 2.0 ** -900, -(2.0 ** -899) * 1.7, 2.0 ** 899 * 1.3]
        for a in avals:
            yield float(a), float(b), rb


def test_rn_div_equals_ieee_division():
    rng = np.random.default_rng(11)
    n = 0
    for a, b, rb in _cases(rng):
        got = rn_div(a, b, rb)
        want = a / b
        assert got == want and np.signbit(got) == np.signbit(want), (a, b, got, want)
        n += 1
    assert n > 10000


def test_reciprocal_product_alone_is_not_enough():
    """RN(a*rb) alone misses the correctly rounded quotient on a large share of
    operands, so the cases above exercise the corrections."""
    rng = np.random.default_rng(12)
    bad = 0
    for _ in range(20000):
        b = float(rng.uniform(1.0, 2.0))
        a = float(rng.uniform(1.0, 2.0))
        rb = 1.0 / b
        q = a * rb
        if q != a / b:
            bad += 1
    assert bad > 0
