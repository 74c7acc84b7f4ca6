"""Pins of the oracle's scalar kernels (CPU).

* I0, I2, I4 against their integral definition
  I_{2k}(lambda) = int_{|lambda|}^inf (s^2 - lambda^2)^k erfc(s) ds (mpmath, 40 digits);
* c(lambda = 0) = (22/5, 8/5, 2/15)/sqrt(pi), which makes the general
  regularized Stokeslet coincide with the independently written on-surface one;
* the r -> 0 self limits and the continuity at r/delta = 8;
* agreement with the compiled reference's values (golden fixture).
"""
import math

import mpmath as mp
import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("lam", [0.0, 0.05, 0.4, -0.9, 1.7, -3.2, 5.0])
def test_In_integral_definition(oracle, lam):
    mp.mp.dps = 40
    got = oracle.In_024(lam)
    L = mp.mpf(lam)
    for k in range(3):
        ref = mp.quad(lambda s: (s * s - L * L) ** k * mp.erfc(s), [abs(L), abs(L) + 4, mp.inf])
        # the closed form cancels for large |lambda|; absolute error relative to the e^{-lambda^2} scale
        scale = float(mp.e ** (-L * L))
        assert abs(got[k] - float(ref)) <= 1e-13 * max(scale, 1e-300) + 1e-14 * abs(float(ref)), (k, got[k], ref)


def test_coefficients_at_lambda_zero(oracle):
    c = oracle.reg_coeff(0.0) * math.sqrt(math.pi)
    assert np.allclose(c, [22 / 5, 8 / 5, 2 / 15], rtol=1e-15, atol=1e-15)


def test_sl_general_equals_onsurface_at_lambda_zero(oracle):
    d = 0.13
    # s2 = (erf + P2 rho e^{-rho^2})/r^3 cancels like 1/rho^2 as r -> 0, so the
    # comparison starts at rho = 0.05
    r = np.linspace(0.05 * d, 8 * d, 500)
    c0 = oracle.reg_coeff(0.0)
    a = oracle.kernel("sl", r, d, c0)
    b = oracle.kernel("sl_on", r, d)
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-12


def test_self_limits(oracle):
    """The r < 1e-14 values are the r -> 0 limits of the regularized kernels:
    the distance to the limit shrinks like rho^2 (ratio 4 per halving)."""
    d = 0.2
    rp = math.sqrt(math.pi)
    lim = {("sl_on", 0): 32 / (5 * rp * d), ("sl_on", 1): 256 / (15 * rp * d ** 3),
           ("dl_on", 1): 128 / (45 * rp * d ** 5)}
    for (kind, comp), L in lim.items():
        z = oracle.kernel(kind, [0.0], d)[0][comp]
        assert abs(z - L) <= 1e-15 * L
        rho = np.array([0.04, 0.02, 0.01])
        v = oracle.kernel(kind, rho * d, d)[:, comp]
        e = np.abs(v - L)
        ratios = e[:-1] / e[1:]
        assert np.all((ratios > 3.6) & (ratios < 4.4)), (kind, comp, ratios)
        assert e[-1] < 1e-3 * L


@pytest.mark.parametrize("kind", ["sl", "sl_on", "dl", "dl_on"])
def test_continuity_at_8_delta(oracle, kind):
    d = 0.07
    c = oracle.reg_coeff(0.3)
    lo, hi = 8 * d * (1 - 1e-13), 8 * d * (1 + 1e-13)
    a, b = oracle.kernel(kind, [lo, hi], d, c)
    assert np.all(np.abs(a - b) <= 1e-11 * np.abs(b))


def test_kernels_match_reference_fixture(oracle):
    g = golden("kernels.npz")
    coef = np.array([oracle.reg_coeff(l) for l in g["lam"]])
    assert np.array_equal(coef, g["coef"])
    for kind in ("sl", "sl_on", "dl", "dl_on"):
        assert np.array_equal(oracle.kernel(kind, g["r"], float(g["d"]), g["c"]), g[kind]), kind
