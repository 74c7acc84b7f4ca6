"""Generates dl_on_table.h: the on-surface stresslet kernel of the 7th-order
method (Stokes_DL_7ord_onSurf, highreg_stokes_2surf/utilities.cpp:291-323) as a
piecewise polynomial of s = rho^2 = (r/delta)^2 on [0, 64].

The kernel is H3 = F(rho) / r^5 with
    F(rho) = erf(rho) + (2/sqrt(pi)) (-1 - 2/3 s + 52/45 s^2 - 8/45 s^3) rho e^{-s};
F is odd with a zero of order 5 at 0, so G(s) = F(rho) / rho^5 is an entire
function of s and H3 = G(s) / delta^5.  The table holds G on NINT intervals of
width 64/NINT, each as a degree-DEG polynomial in the local variable
x = 2 (s - s_mid) / width in [-1, 1]: interpolation at the DEG+1 Chebyshev
points of the first kind in 50-digit arithmetic (mpmath; near s = 0 the
cancellation in F is absorbed by raising the working precision), re-expanded in
monomials of x and rounded to FP64.  The script reports the maximum relative
error of the FP64 Horner evaluation against mpmath on a dense sample.

    python gen_dl_on_table.py > dl_on_table.h
"""
import sys

import mpmath as mp
import numpy as np

NINT = 128
DEG = 10
SMAX = 64


def G(s):
    """F(rho)/rho^5 at s = rho^2 > 0; the working precision grows as s falls,
    since F ~ rho^5 cancels from terms ~ rho."""
    s = mp.mpf(s)
    with mp.workdps(mp.mp.dps + max(0, int(-2.5 * mp.log10(s))) + 10):
        rho = mp.sqrt(s)
        k = 2 / mp.sqrt(mp.pi)
        p3 = -1 - mp.mpf(2) / 3 * s + mp.mpf(52) / 45 * s ** 2 - mp.mpf(8) / 45 * s ** 3
        F = mp.erf(rho) + k * p3 * rho * mp.exp(-s)
        return +(F / rho ** 5)


def main():
    mp.mp.dps = 50
    w = mp.mpf(SMAX) / NINT
    coefs = []
    for i in range(NINT):
        a = i * w
        mid = a + w / 2
        xs = [mp.cos(mp.pi * (2 * k + 1) / (2 * (DEG + 1))) for k in range(DEG + 1)]
        vals = [G(mid + w / 2 * x) for x in xs]
        A = mp.matrix(DEG + 1, DEG + 1)
        for r, x in enumerate(xs):
            for j in range(DEG + 1):
                A[r, j] = x ** j
        c = mp.lu_solve(A, mp.matrix(vals))
        coefs.append([float(c[j]) for j in range(DEG + 1)])
    C = np.array(coefs)
    # FP64 evaluation as the kernels do it: k = int(s / w), x = s * (2/w) - (2k + 1)
    rng = np.random.default_rng(1)
    ss = np.concatenate([rng.uniform(0, SMAX, 3000), np.geomspace(1e-12, 1, 300), [0.0, 63.999999]])
    inv = 2.0 / float(w)
    worst = 0.0
    for s in ss:
        k = min(int(s * (NINT / SMAX)), NINT - 1)
        x = s * inv - (2 * k + 1)
        g = C[k, DEG]
        for j in range(DEG - 1, -1, -1):
            g = g * x + C[k, j]
        ref = G(s) if s > 0 else G(mp.mpf(10) ** -40)
        worst = max(worst, abs(g - float(ref)) / abs(float(ref)))
    out = sys.stdout
    out.write("// Generated by gen_dl_on_table.py: G(s) = F(rho)/rho^5, s = rho^2 in [0, 64], of the\n")
    out.write("// 7th-order on-surface stresslet kernel H3 = F(rho)/r^5 = G(s)/delta^5, piecewise\n")
    out.write(f"// polynomial: {NINT} intervals of width {SMAX}/{NINT}, degree {DEG} in x = 2(s - s_mid)/width.\n")
    out.write(f"// Max relative error of the FP64 evaluation against 50-digit arithmetic: {worst:.2e}.\n")
    out.write("#pragma once\n\nnamespace hor {\n\n")
    out.write(f"constexpr int kDlOnIntervals = {NINT};\n")
    out.write(f"constexpr int kDlOnDeg = {DEG};\n")
    out.write(f"constexpr double kDlOnSmax = {SMAX}.0;\n")
    out.write("// kDlOnTable[k * (kDlOnDeg + 1) + j]: coefficient of x^j on interval k.\n")
    out.write(f"constexpr double kDlOnTable[{NINT * (DEG + 1)}] = {{\n")
    for row in coefs:
        out.write("    " + ", ".join(repr(v) for v in row) + ",\n")
    out.write("};\n\n}  // namespace hor\n")
    print(f"max relative error {worst:.2e}", file=sys.stderr)


if __name__ == "__main__":
    main()
