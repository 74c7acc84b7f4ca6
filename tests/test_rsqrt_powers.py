"""The direct r^-3 and r^-5 of csrc/device.h (rsqrt3_64, rsqrt5_64): from the
hardware estimate y of x^(-1/2) (relative error below 5.2e-8 on gfx950),
e = 1 - x y^2 and x^(-k/2) = y^k (1 + c1 e + c2 e^2) with the binomial
coefficients of (1 - e)^(-k/2).  The same operation sequence is run here
with every fused multiply-add rounded once (exact rational arithmetic) and
compared with 40-digit powers: a wrong coefficient or a dropped term costs
far more than the few ulps allowed.  The device code itself is covered by the
GPU parity tests of the kernels that call it."""
import math
import random
from fractions import Fraction

from mpmath import mp, mpf


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _mul(a, b):
    return float(Fraction(a) * Fraction(b))


def _rsqrt_pow(x, y, k, c1, c2):
    y2 = _mul(y, y)
    e = _fma(-x, y2, 1.0)
    yk = _mul(y2, y) if k == 3 else _mul(_mul(y2, y2), y)
    return _fma(_mul(yk, e), _fma(e, c2, c1), yk)


def test_direct_inverse_powers_within_three_ulps():
    mp.dps = 40
    rng = random.Random(7)
    worst = {3: 0.0, 5: 0.0}
    for _ in range(3000):
        x = math.ldexp(rng.uniform(1.0, 2.0), rng.randint(-40, 40))
        exact = 1 / mp.sqrt(mpf(x))
        y = float(exact * (1 + mpf(rng.uniform(-5.2e-8, 5.2e-8))))
        for k, c1, c2 in ((3, 1.5, 1.875), (5, 2.5, 4.375)):
            v = _rsqrt_pow(x, y, k, c1, c2)
            ex = exact ** k
            worst[k] = max(worst[k], abs(float((mpf(v) - ex) / ex)))
    ulp = 2.0 ** -53
    print({k: f"{v / ulp:.2f} ulp" for k, v in worst.items()})
    assert worst[3] <= 3 * ulp
    assert worst[5] <= 3.5 * ulp
