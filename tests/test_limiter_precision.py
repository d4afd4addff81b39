"""The FCT limiter's positivity test in the FP32 form is exact.

The limiter (FCT_DeviceMath.cuh:108-125, `fct_limit` in khs_device.h) maps
an antidiffusive flux f and three neighbour differences of the low-order
solution to

    s = sign(d0),  L = s * max(0, min(s * dm1, |f|, s * dp1)).

In the FP32 form f and the operands of the differences are FP32 values and
each difference is one FP32 subtraction.  IEEE round-to-nearest is monotone
and odd and keeps the sign of a difference of two floats (a - b rounds to 0
only when a == b), and min, max, |.| and the multiplication by s = +-1 are
exact.  So the FP32 evaluation equals the limiter evaluated in exact
arithmetic on the same FP32 operands and rounded once to FP32:

    fl32(L_exact) == L_fp32   bit for bit.

An FP64 positivity test (differences, sign and clipping in FP64, the result
rounded to FP32 for the update) can only equal that or, where an FP64
difference is itself rounded (operands more than 2^29 apart in magnitude)
and lands on an FP32 rounding midpoint, differ from it by double rounding.
These tests check the identity on random operands over wide magnitude
ranges, with exact rational arithmetic where FP64 differences would not be
exact."""
from fractions import Fraction

import numpy as np

f32 = np.float32


def kmin(a, b):
    """The kernels' select min: b if b < a else a (fixes the sign of a zero result)."""
    return np.where(b < a, b, a)


def kmax(a, b):
    return np.where(a < b, b, a)


def limit(f, dm1, d0, dp1):
    """The limiter in numpy's dtype of its inputs (elementwise), with the
    kernels' select order: kmax(0, kmin(s dm1, kmin(|f|, s dp1)))."""
    one = np.ones_like(f)
    s = np.copysign(one, d0)
    t = kmin(s * dm1, kmin(np.abs(f), s * dp1))
    return s * kmax(np.zeros_like(f), t)


def round_f32(x):
    """Correctly rounded FP32 value of the rational x (ties to even)."""
    c = f32(float(x))
    best = None
    for cand in (np.nextafter(c, f32(-np.inf)), c, np.nextafter(c, f32(np.inf))):
        err = abs(Fraction(float(cand)) - x)
        if best is None or err < best[0] or (err == best[0] and (cand.view(np.uint32) & 1) == 0):
            best = (err, cand)
    return best[1]


def limit_exact(f, l0, l1, l2, l3):
    """fl32 of the limiter evaluated exactly on FP32 operands."""
    F = Fraction(float(f))
    L = [Fraction(float(v)) for v in (l0, l1, l2, l3)]
    dm1, d0, dp1 = L[1] - L[0], L[2] - L[1], L[3] - L[2]
    s = -1 if d0 < 0 else 1
    lim = max(Fraction(0), min(s * dm1, abs(F), s * dp1))
    if lim == 0:
        return f32(s * 0.0)                       # IEEE s * (+0): the zero carries the sign of d0
    return round_f32(s * lim)


def operands(rng, n, span):
    """FP32 low-order values l0..l3 and fluxes with magnitudes over 2^-span..2^span,
    random signs, and a share of equal neighbours and zeros."""
    def draw():
        m = rng.uniform(0.5, 1.0, n) * np.exp2(rng.integers(-span, span + 1, n)).astype(np.float64)
        v = (m * rng.choice([-1.0, 1.0], n)).astype(f32)
        v[rng.random(n) < 0.05] = f32(0.0)
        return v
    ls = [draw() for _ in range(4)]
    for k in range(1, 4):                         # equal neighbours: zero differences
        eq = rng.random(n) < 0.05
        ls[k][eq] = ls[k - 1][eq]
    return draw(), ls


def test_fp32_limiter_equals_exact_limiter_rounded_once():
    rng = np.random.default_rng(7)
    n = 400_000
    f, (l0, l1, l2, l3) = operands(rng, n, span=12)
    # magnitudes within 2^25 of each other: FP64 differences are exact, so the
    # FP64 evaluation is the exact one
    r32 = limit(f, l1 - l0, l2 - l1, l3 - l2)
    d = [np.float64(x) for x in (l0, l1, l2, l3)]
    r64 = limit(f.astype(np.float64), d[1] - d[0], d[2] - d[1], d[3] - d[2]).astype(f32)
    assert r32.dtype == f32
    np.testing.assert_array_equal(r32.view(np.uint32), r64.view(np.uint32))
    # the limited flux never exceeds |f| and has the sign of d0 (or is zero)
    assert np.all(np.abs(r32) <= np.abs(f))
    s = np.copysign(np.ones_like(f), l2 - l1)
    assert np.all(r32 * s >= 0)


def test_fp32_limiter_equals_exact_limiter_at_large_magnitude_gaps():
    rng = np.random.default_rng(11)
    n = 3000
    f, (l0, l1, l2, l3) = operands(rng, n, span=60)   # gaps up to 2^120: FP64 differences round
    r32 = limit(f, l1 - l0, l2 - l1, l3 - l2)
    ref = np.array([limit_exact(*args) for args in zip(f, l0, l1, l2, l3)], dtype=f32)
    np.testing.assert_array_equal(r32.view(np.uint32), ref.view(np.uint32))


def test_round_f32_is_correct_rounding():
    # midpoint between 1 and the next float rounds to even (1); just above it rounds up
    one = Fraction(1)
    ulp = Fraction(1, 2 ** 23)
    assert round_f32(one + ulp / 2) == f32(1.0)
    assert round_f32(one + ulp / 2 + Fraction(1, 2 ** 60)) == np.nextafter(f32(1.0), f32(2.0))
    assert round_f32(Fraction(-3, 7)) == f32(-3.0 / 7.0)
