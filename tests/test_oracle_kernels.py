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


def test_on_surface_stresslet_table_matches_the_kernel(oracle):
    """The on-surface stresslet kernel the GPU reads from its table
    (csrc/dl_on_table.h: G(s) = F(rho)/rho^5, H3 = G/delta^5):
    * against F(rho)/r^5 in 60-digit arithmetic (mpmath) on rho in (0, 8]:
      2e-15 relative (a few ulps);
    * against the oracle's Stokes_DL_7ord_onSurf (utilities.cpp:291-323) for
      rho >= 0.5, 1e-13 relative; below that the FP64 formula itself loses
      digits to cancellation (F ~ rho^5 from terms ~ rho), e.g. 2e-11 at
      rho = 0.05, which the table does not."""
    import pathlib
    import re
    import mpmath as mp
    src = (pathlib.Path(__file__).parents[1] / "high-order-regularization_amd" / "csrc" / "dl_on_table.h").read_text()
    nint = int(re.search(r"kDlOnIntervals = (\d+)", src).group(1))
    deg = int(re.search(r"kDlOnDeg = (\d+)", src).group(1))
    smax = float(re.search(r"kDlOnSmax = ([0-9.]+)", src).group(1))
    body = src[src.index("kDlOnTable["):]
    vals = [float(v) for v in re.findall(r"-?\d\.\d+(?:e[-+]\d+)?", body[body.index("{"):body.index("}")])]
    C = np.array(vals).reshape(nint, deg + 1)

    def table_H3(r, d):
        s = (r / d) ** 2
        k = np.minimum((s * (nint / smax)).astype(int), nint - 1)
        x = s * (2 * nint / smax) - (2 * k + 1)
        g = C[k, deg]
        for j in range(deg - 1, -1, -1):
            g = g * x + C[k, j]
        return g / d ** 5

    d = 0.1
    mp.mp.dps = 60
    rho = np.concatenate([np.geomspace(1e-4, 1, 60), np.linspace(1, 8, 240)])
    exact = []
    for p in rho:
        q = mp.mpf(p)
        k = 2 / mp.sqrt(mp.pi)
        p3 = -1 - mp.mpf(2) / 3 * q ** 2 + mp.mpf(52) / 45 * q ** 4 - mp.mpf(8) / 45 * q ** 6
        F = mp.erf(q) + k * p3 * q * mp.exp(-q * q)
        exact.append(float(F / (q * d) ** 5))
    exact = np.array(exact)
    assert np.abs(table_H3(rho * d, d) / exact - 1).max() < 2e-15   # a few ulps, the /delta^5 included
    r = np.linspace(0.5 * d, 8 * d, 600)
    ref = oracle.kernel("dl_on", r, d)[:, 1]
    assert np.abs(table_H3(r, d) / ref - 1).max() < 1e-13


def test_erf_exp_table_matches_erf_and_exp():
    """erf(rho)/rho and exp(-rho^2) as the GPU reads them from its table
    (csrc/erf_exp_table.h, piecewise polynomials of s = rho^2 on [0, 64])
    against 60-digit arithmetic: 1 ulp-level (4e-16 relative); and rho A(s)
    against the C library's erf, exp(-s) against its exp (the functions the
    reference calls): 1e-15."""
    import math
    import pathlib
    import re
    import mpmath as mp
    src = (pathlib.Path(__file__).parents[1] / "high-order-regularization_amd" / "csrc" / "erf_exp_table.h").read_text()
    nint = int(re.search(r"kAEIntervals = (\d+)", src).group(1))
    deg = int(re.search(r"kAEDeg = (\d+)", src).group(1))
    smax = float(re.search(r"kAESmax = ([0-9.]+)", src).group(1))
    body = src[src.index("kErfExpTable["):]
    vals = [float(v) for v in re.findall(r"-?\d\.\d+(?:e[-+]\d+)?", body[body.index("{"):body.index("}")])]
    C = np.array(vals).reshape(nint, 2, deg + 1)
    rng = np.random.default_rng(3)
    s = np.concatenate([rng.uniform(0, smax, 500), np.geomspace(1e-10, 1, 50), [0.0, smax]])
    k = np.minimum((s * (nint / smax)).astype(int), nint - 1)
    x = s * (2 * nint / smax) - (2 * k + 1)
    out = []
    for f in range(2):
        g = C[k, f, deg]
        for j in range(deg - 1, -1, -1):
            g = g * x + C[k, f, j]
        out.append(g)
    A, E = out
    mp.mp.dps = 60
    A_ex = np.array([float(mp.erf(mp.sqrt(v)) / mp.sqrt(v)) if v > 0 else float(2 / mp.sqrt(mp.pi)) for v in s])
    E_ex = np.array([float(mp.exp(-mp.mpf(v))) for v in s])
    assert np.abs(A / A_ex - 1).max() < 4e-16
    assert np.abs(E / E_ex - 1).max() < 4e-16
    rho = np.sqrt(s)
    libc_erf = np.array([math.erf(r) for r in rho])
    nz = rho > 0
    assert np.abs(rho[nz] * A[nz] / libc_erf[nz] - 1).max() < 1e-15
    assert np.abs(E / np.exp(-s) - 1).max() < 1e-15
