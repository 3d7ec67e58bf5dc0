"""The vertical-divisor division of the stencil kernels (zdiv in les_common.h): FP32 a / b from
the FP64 pair (b, y = RN64(1/b)) as RN32(q0 + (a - b q0) y) with q0 = RN64(a y), each FMA
rounded once.  Restated here in exact rational arithmetic (an FMA is the exact value rounded
to the nearest double) and checked against numpy's IEEE FP32 division for the benchmark grid's
divisors (dzn, dzs and neighbour sums of a 2 m / x1.05 stretched grid), random numerators over
the whole exponent range, subnormal numerators, and quotients that are exactly midpoints
between subnormal floats, where the uncorrected q0 must fail somewhere (so the midpoint cases
exercise the correction).  Zero, infinite and NaN operands go through the hardware fixup
instruction (the division's own last step) and are not restated here."""
from fractions import Fraction

import numpy as np

f32 = np.float32


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _zdiv(a, b):
    x = float(a)
    y = 1.0 / float(b)
    q0 = x * y
    r = _fma(-float(b), q0, x)
    return f32(_fma(r, y, q0)), f32(q0)


def _grid_divisors(kp=90):
    dzn = [f32(2.0)] * 21
    while len(dzn) < kp + 4:
        dzn.append(f32(dzn[-1] * f32(1.05)))
    out = set()
    for k in range(len(dzn) - 1):
        out.add(float(dzn[k]))
        out.add(float(f32(dzn[k + 1] / f32(2)) + f32(dzn[k] / f32(2))))
        out.add(float(f32(dzn[k] + dzn[k + 1])))
    return sorted(out)


def _same(x, y):
    return np.asarray(x, f32).view(np.uint32) == np.asarray(y, f32).view(np.uint32)


def test_zdiv_equals_division_on_grid_divisors():
    rng = np.random.default_rng(20240611)
    divs = _grid_divisors()
    bits = rng.integers(0, 2 ** 32, size=60, dtype=np.uint64).astype(np.uint32)
    nums = [f32(v) for v in bits.view(np.float32) if np.isfinite(v) and v != 0]
    nums += [f32(v) for v in rng.standard_normal(40).astype(np.float32)]
    nums += [np.uint32(m).view(np.float32) for m in (1, 3, 0x7fffff, 0x400001, 0x12345)]
    checked = 0
    for b in divs:
        bf = f32(b)
        for a in nums:
            got, _ = _zdiv(a, bf)
            assert _same(got, f32(a) / bf), (a, b)
            checked += 1
    assert checked > 20000


def test_zdiv_breaks_exact_subnormal_ties_to_even():
    # b = B 2^t with odd B and t >= 1, a = K B 2^(t-150) with odd K and K B < 2^24: a is a
    # float and a / b is exactly the midpoint K 2^-150 between two subnormals.  RN64(a y) can
    # miss the midpoint when K's significand is near 2 and y's rounding error near its bound
    # (b = 4042, 4046 with K close to 2^24 / B).
    ties = 0
    uncorrected_wrong = 0
    for b in _grid_divisors() + [6.0, 10.0, 40.5 * 2, 4042.0, 4046.0]:
        bf = f32(b)
        B = Fraction(float(bf)).numerator
        while B % 2 == 0:
            B //= 2
        kmax = (1 << 24) // B
        for K in list(range(1, 200, 2)) + list(range(max(1, kmax - 130) | 1, kmax, 2)):
            exact = Fraction(K) * Fraction(float(bf)) / 2 ** 150
            a = f32(float(exact))
            if a == 0 or Fraction(float(a)) != exact:
                continue
            want = a / bf
            got, plain = _zdiv(a, bf)
            assert _same(got, want), (float(a), b, K)
            ties += 1
            uncorrected_wrong += int(not _same(plain, want))
    assert ties > 500
    assert uncorrected_wrong > 0  # the correction step is what gets these right
