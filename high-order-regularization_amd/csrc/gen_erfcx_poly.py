"""Generates erfcx_poly.h: the polynomial R(t) ~ erfcx(rho) = erfc(rho) exp(rho^2)
in t = 1/(1 + rho/3), for rho in [0, 8], used by the sum kernels as
erf(rho) = 1 - exp(-rho^2) R(t).

Method: interpolation of erfcx at the 19 Chebyshev points of the first kind on
t in [3/11, 1] (rho = 8 .. 0), evaluated in 50-digit arithmetic (mpmath), then
re-expanded in monomials of u = t - t_mid and rounded to FP64; Horner's rule
in u is well conditioned there (max |c_k| < 2).  The script also reports the
maximum absolute error of the resulting erf in FP64 on a dense grid of rho.

    python gen_erfcx_poly.py > erfcx_poly.h
"""
import sys

import mpmath as mp
import numpy as np

DEG = 18
KAPPA = 3


def main():
    mp.mp.dps = 50
    kappa = mp.mpf(KAPPA)
    a, b = 1 / (1 + mp.mpf(8) / kappa), mp.mpf(1)
    mid = (a + b) / 2
    nodes = [mid + (b - a) / 2 * mp.cos(mp.pi * (2 * k + 1) / (2 * (DEG + 1))) for k in range(DEG + 1)]
    vals = [mp.erfc(kappa * (1 / t - 1)) * mp.exp((kappa * (1 / t - 1)) ** 2) for t in nodes]
    A = mp.matrix(DEG + 1, DEG + 1)
    for i, t in enumerate(nodes):
        for j in range(DEG + 1):
            A[i, j] = (t - mid) ** j
    c = mp.lu_solve(A, mp.matrix(vals))
    cd = [float(c[j]) for j in range(DEG + 1)]
    # accuracy of erf = 1 - exp(-rho^2) R in FP64
    rho = np.concatenate([np.linspace(0, 8, 4001), np.geomspace(1e-6, 8, 1000)])
    u = 1 / (1 + rho / KAPPA) - float(mid)
    R = np.zeros_like(u)
    for ck in reversed(cd):
        R = R * u + ck
    err = max(abs(1 - np.exp(-r * r) * Rv - float(mp.erf(mp.mpf(r)))) for r, Rv in zip(rho, R))
    out = sys.stdout
    out.write("// Generated by gen_erfcx_poly.py: R(t) ~ erfcx(rho) = erfc(rho) exp(rho^2),\n")
    out.write(f"// t = 1/(1 + rho/{KAPPA}), rho in [0, 8]; R = sum_k c_k (t - t_mid)^k.\n")
    out.write(f"// Max |1 - exp(-rho^2) R - erf(rho)| in FP64 over rho in [0, 8]: {err:.2e}.\n")
    out.write("#pragma once\n\nnamespace hor {\n\n")
    out.write(f"constexpr int kErfcxDeg = {DEG};\n")
    out.write(f"constexpr double kErfcxKappaInv = 1.0 / {KAPPA}.0;\n")
    out.write(f"constexpr double kErfcxTmid = {float(mid)!r};\n")
    out.write("// The plans copy these to device memory; the kernels read them through a\n")
    out.write("// pointer as wave-uniform scalar loads, so that each Horner step is one FMA\n")
    out.write("// with an SGPR operand.\n")
    out.write("constexpr double kErfcxC[kErfcxDeg + 1] = {\n")
    for v in cd:
        out.write(f"    {v!r},\n")
    out.write("};\n\n}  // namespace hor\n")
    print(f"max abs erf error {err:.2e}", file=sys.stderr)


if __name__ == "__main__":
    main()
