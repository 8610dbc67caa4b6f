"""The lean step kernels divide by the grid spacings with a four-operation
sequence instead of the IEEE division instruction sequence:

    zh = RD(1/d), zl = RN(1/d - zh), y = RN(1/d)
    q0 = RN(x*zl + RN(x*zh));  r = RN(q0*d - x);  q1 = RN(q0 - r*y)

claimed to return RN(x/d) exactly (and the sign of a zero x) for d in
[2^-40, 2^40] and x = 0 or 2^-960 <= |x| <= 2^980.  These CPU tests pin the
claim: exhaustively in a 11-bit-significand format (every x and d
significand, exact integer arithmetic), and in binary64 on random and
near-halfway numerators for the spacings the tests and benchmarks use, with
every fused multiply-add evaluated exactly in rationals.  The GPU parity
tests then compare the device kernels with the oracle bit for bit.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest


# ------------------------------------------------------------ binary16 model
def _rn16(v):
    """Nearest binary16 of exact float64 values (ties to even)."""
    return np.asarray(v, dtype=np.float64).astype(np.float16).astype(np.float64)


def _rn_int_quotient(X, D, bits):
    """RN of X / D (integers, quotient in [1, 2) or [0.5, 1)) to `bits`
    significand bits, as an exact float64."""
    X = X.astype(np.int64)
    hi = X >= D
    shift = np.where(hi, bits - 1, bits)
    num = X << shift
    q, rem = num // D, num % D
    up = (2 * rem > D) | ((2 * rem == D) & (q % 2 == 1))
    q = q + up
    return q.astype(np.float64) * np.exp2(-shift.astype(np.float64))


def _recip16(D):
    """(RN, RD) of 1/d for d = D / 1024 in binary16, exactly."""
    inv = Fraction(1024, D)
    c = float(np.float16(float(inv)))
    cands = sorted({c, float(np.nextafter(np.float16(c), np.float16(np.inf))),
                    float(np.nextafter(np.float16(c), np.float16(-np.inf)))})
    rn = min(cands, key=lambda v: (abs(Fraction(v) - inv),
                                   int(np.float16(v).view(np.uint16)) & 1))
    rd = max(v for v in cands if Fraction(v) <= inv)
    return rn, rd


def test_fast_division_exhaustive_binary16():
    X = np.arange(1024, 2048)
    x = X / 1024.0
    bad = 0
    for D in range(1024, 2048):
        d = D / 1024.0
        y, zh = _recip16(D)
        zl = float(np.float16(float(Fraction(1024, D) - Fraction(zh))))
        assert zl >= 0.0
        q0 = _rn16(x * zl + _rn16(x * zh))  # both sums exact in float64
        r = _rn16(q0 * d - x)
        q1 = _rn16(q0 - r * y)
        bad += int(np.count_nonzero(q1 != _rn_int_quotient(X, D, 11)))
    assert bad == 0


# ------------------------------------------------------------ binary64
def _fma(a, b, c):
    """IEEE fused multiply-add, round to nearest, signed zeros included."""
    v = Fraction(a) * Fraction(b) + Fraction(c)
    if v != 0:
        return float(v)
    prod_zero = a == 0 or b == 0
    prod_neg = (math.copysign(1, a) * math.copysign(1, b)) < 0
    if prod_zero and c == 0 and prod_neg and math.copysign(1, c) < 0:
        return -0.0
    return 0.0


def _constants(d):
    """As fast_division_constants in the kernels' host code."""
    y = 1.0 / d
    if _fma(-y, d, 1.0) < 0.0:
        y = math.nextafter(y, 0.0)
    zl = _fma(-y, d, 1.0) / d
    return y, zl


def _fast_div(x, d, zh, zl, y):
    q0 = _fma(x, zl, x * zh)
    r = _fma(q0, d, -x)
    return _fma(-r, y, q0)


@pytest.mark.parametrize("d", [1.249, 1.25, 1.1, 5.0, 0.8, 3.0, 1e-3, 7777.7,
                               2.0 ** -40, 2.0 ** 40 * 0.999])
def test_fast_division_binary64(d):
    zh, zl = _constants(d)
    y = 1.0 / d
    assert zl >= 0.0 and Fraction(zh) <= 1 / Fraction(d)
    rng = random.Random(hash(d) & 0xFFFF)
    xs = [0.0, -0.0, 2.0 ** -960, -(2.0 ** 980), d, -d, 3 * d, d * (1 + 2 ** -52)]
    for _ in range(3000):
        m = 1.0 + rng.random()
        xs.append(rng.choice((-1, 1)) * math.ldexp(m, rng.randint(-959, 979)))
    for _ in range(3000):  # numerators whose quotient is near a halfway point
        q = math.ldexp(1.0 + rng.getrandbits(52) / 2.0 ** 52, rng.randint(-200, 200))
        half = Fraction(q) + Fraction(math.ulp(q)) / 2
        xs.append(float(half * Fraction(d)))
    for x in xs:
        got = _fast_div(x, d, zh, zl, y)
        want = x / d
        assert got == want and math.copysign(1, got) == math.copysign(1, want), (x, d)
