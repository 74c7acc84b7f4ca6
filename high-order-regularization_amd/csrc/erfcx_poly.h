// Generated by gen_erfcx_poly.py: R(t) ~ erfcx(rho) = erfc(rho) exp(rho^2),
// t = 1/(1 + rho/3), rho in [0, 8]; R = sum_k c_k (t - t_mid)^k.
// Max |1 - exp(-rho^2) R - erf(rho)| in FP64 over rho in [0, 8]: 6.72e-16.
#pragma once

namespace hor {

constexpr int kErfcxDeg = 18;
constexpr double kErfcxKappaInv = 1.0 / 3.0;
constexpr double kErfcxTmid = 0.6363636363636364;
// The plans copy these to device memory; the kernels read them through a
// pointer as wave-uniform scalar loads, so that each Horner step is one FMA
// with an SGPR operand.
constexpr double kErfcxC[kErfcxDeg + 1] = {
    0.28972211632346423,
    1.0004442745763151,
    1.6227282288500673,
    1.9838713379588286,
    1.7432797096177521,
    0.9468511766696378,
    0.10735803912002746,
    -0.25122689808545384,
    -0.12159808205797053,
    0.07277983270101679,
    0.06007265664418884,
    -0.03197951821885252,
    -0.027078970642441223,
    0.019952569062346773,
    0.010340830745889616,
    -0.01329967104159605,
    -0.0017934405906309446,
    0.006394934421347902,
    -0.0010558403020711636,
};

}  // namespace hor
