"""Division by a launch-wide constant in the FP64 kernels (CDiv in
khs_device.h): q0 = RN(a y) with y = RN(1/b), then two corrections
q = RN(q - RN(b q - a) y), each an fma, must give the bits of IEEE a / b
(Markstein's theorem: after the first correction q is within one ulp of a / b,
and with y within half an ulp of 1/b the second rounds a / b correctly).
Here fma is evaluated exactly with rationals and rounded once, and the result
is compared with Python's IEEE division for random operands over a wide range
of magnitudes and for the grid spacings the workloads use."""
import random
from fractions import Fraction

import pytest


def fma(a, b, c):
    """RN(a b + c) for nonzero exact results (the operands here give no zero)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def cdiv(a, b, y):
    q = a * y
    rn = fma(b, q, -a)
    q = fma(-rn, y, q)
    rn = fma(b, q, -a)
    return fma(-rn, y, q)


def one_correction(a, b, y):
    q = a * y
    return fma(-fma(b, q, -a), y, q)


@pytest.mark.parametrize("b", [0.1, 0.05, 0.025, 1.0, 0.2, 0.3, 1.0 / 3.0, 6.4 / 64, 0.0987654321, 25.6 / 256])
def test_corrected_quotient_is_ieee_division(b):
    rnd = random.Random(int(b * 1e9))
    y = 1.0 / b
    for _ in range(3000):
        a = rnd.uniform(1.0, 2.0) * 2.0 ** rnd.randint(-60, 60) * rnd.choice((-1.0, 1.0))
        assert cdiv(a, b, y) == a / b, (a, b)
    # exact quotients (zero remainder) and neighbours of them
    for k in range(1, 200):
        a = k * b
        for t in (a, a * (1 + 2 ** -52), a * (1 - 2 ** -53)):
            assert cdiv(t, b, y) == t / b


def test_first_correction_is_within_one_ulp():
    """The theorem's premise for the second correction: after the first, q is
    within one ulp of a / b."""
    rnd = random.Random(5)
    b = 0.1
    y = 1.0 / b
    for _ in range(3000):
        a = rnd.uniform(1.0, 2.0) * 2.0 ** rnd.randint(-40, 40)
        q1, e = one_correction(a, b, y), a / b
        assert abs(q1 - e) <= 2.0 ** -52 * abs(e)
