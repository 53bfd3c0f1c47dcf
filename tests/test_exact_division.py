"""Division by a mesh constant without a divide (csrc/fast.h Recip, csrc/launch.h
by_const): the FP64 form replaces a / b, for b the uniform mesh spacing or 6,
by a faithful quotient from a double-word reciprocal and one Markstein
correction.  The fused kernels claim the result is a / b bit for bit; this
checks the algorithm against IEEE division on random quotients, on quotients
placed next to rounding midpoints (where a wrong rounding would show), and on
signed zeros.  The FMA is emulated exactly with rationals.
"""
import math
import random
import struct
from fractions import Fraction

import pytest


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _fma(x: float, y: float, z: float) -> float:
    """x*y + z rounded once (IEEE 754 fusedMultiplyAdd, round to nearest)."""
    exact = Fraction(x) * Fraction(y) + Fraction(z)
    if exact != 0:
        return float(exact)  # int / int true division rounds correctly
    # an exact zero sum: +0 unless both the product and z are -0
    prod_neg = (math.copysign(1.0, x) * math.copysign(1.0, y)) < 0
    return -0.0 if (prod_neg and math.copysign(1.0, z) < 0) else 0.0


def _recip(b: float):
    """fast.h recip_of."""
    hn = 1.0 / b
    hd = hn if _fma(-hn, b, 1.0) >= 0.0 else math.nextafter(hn, 0.0)
    ld = _fma(-hd, b, 1.0) / b
    return b, hn, hd, ld


def _by_const(a: float, r) -> float:
    """launch.h by_const, FP64 form."""
    b, hn, hd, ld = r
    q = _fma(a, hd, a * ld)
    e = _fma(q, b, -a)
    return _fma(-e, hn, q)


# mesh spacings of decks of 10 units over n cells (widened from the single
# precision cell count as initialise_chunk.f90 does), the limiter's 6, and
# values whose reciprocal rounds up and down
SPACINGS = [10.0 / float(n) for n in (10, 16, 24, 32, 48, 64, 96, 100, 128, 200, 256, 512, 640, 960)]
SPACINGS += [6.0, 3.0, 0.1, 0.3, 1.0 / 3.0, 0.7, 1.5, 0.0390625]


def test_recip_parts():
    for b in SPACINGS:
        _, hn, hd, ld = _recip(b)
        assert hn == 1.0 / b
        assert ld >= 0.0
        assert Fraction(hd) <= Fraction(1) / Fraction(b)
        # hd + ld is 1/b to far better than one ulp of hd
        err = abs(Fraction(hd) + Fraction(ld) - Fraction(1) / Fraction(b))
        assert err <= Fraction(math.ulp(hd)) / 2**50


@pytest.mark.parametrize("b", SPACINGS)
def test_random_quotients(b):
    rng = random.Random(int(b * 1e6))
    r = _recip(b)
    for _ in range(1500):
        a = math.ldexp(1.0 + rng.random(), rng.randint(-60, 60)) * rng.choice((1.0, -1.0))
        assert _bits(_by_const(a, r)) == _bits(a / b), (a, b)


@pytest.mark.parametrize("b", SPACINGS)
def test_quotients_next_to_midpoints(b):
    """a/b within a few ulps of a of the midpoint between two doubles: the
    cases where a quotient one rounding step off would round the other way."""
    rng = random.Random(7 + int(b * 1e6))
    r = _recip(b)
    for _ in range(800):
        q = math.ldexp(1.0 + rng.random(), rng.randint(-30, 30))
        mid = Fraction(q) + Fraction(math.ulp(q)) / 2
        a = float(Fraction(b) * mid)
        for cand in (a, math.nextafter(a, math.inf), math.nextafter(a, -math.inf)):
            assert _bits(_by_const(cand, r)) == _bits(cand / b), (cand, b)


def test_signed_zero():
    for b in SPACINGS:
        r = _recip(b)
        for a in (0.0, -0.0):
            assert _bits(_by_const(a, r)) == _bits(a / b)


def test_halved_difference_matches_doubled_divisor():
    """viscosity's pressure gradient divides by celldx + celldx; the fused
    kernel halves the numerator instead (both steps exact)."""
    rng = random.Random(3)
    for b in SPACINGS:
        r = _recip(b)
        for _ in range(300):
            p1, p2 = rng.uniform(0.01, 10.0), rng.uniform(0.01, 10.0)
            d = p1 - p2
            assert _bits(_by_const(0.5 * d, r)) == _bits(d / (b + b))
