"""Generates erf_exp_table.h: erf(rho)/rho and exp(-rho^2) as piecewise
polynomials of s = rho^2 on [0, 64], the two special functions of every
regularized kernel off the surface (Stokes_SL_7ord, Stokes_DL_7ord,
Lapl_DL_7ord; highreg_stokes_2surf/utilities.cpp:179-290).

A(s) = erf(sqrt(s))/sqrt(s) and E(s) = exp(-s) are entire functions of s, so a
kernel writes erf(rho) = rho A(s) and rho exp(-rho^2) = rho E(s) and, with
rho / r = 1/delta, needs neither exp nor erf nor a square root for them.  The
table holds both on NINT intervals of width 64/NINT as degree-DEG polynomials in
the local variable x = 2 (s - s_mid)/width in [-1, 1] (interpolation at the
Chebyshev points of the first kind in 50-digit arithmetic, re-expanded in
monomials of x, rounded to FP64), interleaved per interval: DEG+1 coefficients
of A, then DEG+1 of E.  The script reports the maximum relative error of the
FP64 Horner evaluation of each against mpmath.

    python gen_erf_exp_table.py > erf_exp_table.h
"""
import sys

import mpmath as mp
import numpy as np

NINT = 128
DEG = 10
SMAX = 64


def A(s):
    s = mp.mpf(s)
    if s == 0:
        return 2 / mp.sqrt(mp.pi)
    r = mp.sqrt(s)
    return mp.erf(r) / r


def E(s):
    return mp.exp(-mp.mpf(s))


def fit(f, mid, half):
    xs = [mp.cos(mp.pi * (2 * k + 1) / (2 * (DEG + 1))) for k in range(DEG + 1)]
    M = mp.matrix(DEG + 1, DEG + 1)
    for r, x in enumerate(xs):
        for j in range(DEG + 1):
            M[r, j] = x ** j
    c = mp.lu_solve(M, mp.matrix([f(mid + half * x) for x in xs]))
    return [float(c[j]) for j in range(DEG + 1)]


def main():
    mp.mp.dps = 50
    w = mp.mpf(SMAX) / NINT
    rows = []
    for i in range(NINT):
        mid = i * w + w / 2
        rows.append(fit(A, mid, w / 2) + fit(E, mid, w / 2))
    C = np.array(rows)
    rng = np.random.default_rng(1)
    ss = np.concatenate([rng.uniform(0, SMAX, 3000), np.geomspace(1e-12, 1, 200), [0.0, 63.999999]])
    inv = 2.0 / float(w)
    worst = [0.0, 0.0]
    for s in ss:
        k = min(int(s * (NINT / SMAX)), NINT - 1)
        x = s * inv - (2 * k + 1)
        for f, (o, fn) in enumerate(((0, A), (DEG + 1, E))):
            g = C[k, o + DEG]
            for j in range(DEG - 1, -1, -1):
                g = g * x + C[k, o + j]
            ref = float(fn(s))
            worst[f] = max(worst[f], abs(g - ref) / abs(ref))
    out = sys.stdout
    out.write("// Generated by gen_erf_exp_table.py: A(s) = erf(rho)/rho and E(s) = exp(-rho^2),\n")
    out.write(f"// s = rho^2 in [0, {SMAX}], piecewise polynomials: {NINT} intervals, degree {DEG} in\n")
    out.write("// x = 2(s - s_mid)/width.  Max relative error of the FP64 evaluation against\n")
    out.write(f"// 50-digit arithmetic: A {worst[0]:.2e}, E {worst[1]:.2e}.\n")
    out.write("#pragma once\n\nnamespace hor {\n\n")
    out.write(f"constexpr int kAEIntervals = {NINT};\n")
    out.write(f"constexpr int kAEDeg = {DEG};\n")
    out.write(f"constexpr double kAESmax = {SMAX}.0;\n")
    out.write("// kErfExpTable[k * 2 (kAEDeg + 1) + j]: x^j coefficient of A on interval k;\n")
    out.write("// [... + kAEDeg + 1 + j]: of E.\n")
    out.write(f"constexpr double kErfExpTable[{NINT * 2 * (DEG + 1)}] = {{\n")
    for row in rows:
        out.write("    " + ", ".join(repr(v) for v in row) + ",\n")
    out.write("};\n\n}  // namespace hor\n")
    print(f"max relative error A {worst[0]:.2e}, E {worst[1]:.2e}", file=sys.stderr)


if __name__ == "__main__":
    main()
