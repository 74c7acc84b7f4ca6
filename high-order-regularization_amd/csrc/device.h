// Device-side building blocks shared by the sum kernels (HIP, gfx950).
//
// Pair arithmetic follows the reference's per-pair expressions
// (highreg_stokes_2surf/utilities.cpp:179-323, stokeslets.cpp:106-130,
// stresslets.cpp:117-168, 244-269).  The per-target regularization
// polynomial coefficients are precomputed on the host (TgtCoef below).
#pragma once

#include <hip/hip_runtime.h>

#include "dl_on_table.h"
#include "erf_exp_table.h"
#include "hor_internal.h"

#define HOR_HIP_TRY(expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return hor::set_error(HOR_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                           \
  } while (0)

#define HOR_TRY(expr)        \
  do {                       \
    int r_ = (expr);         \
    if (r_ != HOR_OK) return r_; \
  } while (0)

struct hor_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_total_ms = 0, last_kernel_ms = 0;
};

namespace hor {

constexpr int kBlock = 256;     // threads per workgroup (4 waves of 64)
constexpr int kWave = 64;
constexpr double kPi = 3.14159265358979323846;

// Device buffer owned by a plan; freed on destruction.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return set_error(HOR_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                       hipGetErrorString(e));
    }
    n = count;
    return HOR_OK;
  }
  int upload(const T* src, size_t count, hipStream_t s) {
    HOR_TRY(alloc(count));
    if (count) {
      hipError_t e = hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return set_error(HOR_E_HIP, "upload: %s", hipGetErrorString(e));
    }
    return HOR_OK;
  }
  int upload(const std::vector<T>& v, hipStream_t s) { return upload(v.data(), v.size(), s); }
};

// ---------------------------------------------------------------------------
// Per-target regularization data, structure-of-arrays with stride ntp.
//   SL: s1 = (erf(rho) + P1(rho^2) rho e^{-rho^2}) / r,
//       s2 = (erf(rho) + P2(rho^2) rho e^{-rho^2}) / r^3
//       P1 = q0 + q1 rho^2 + q2 rho^4, P2 = q3 + q4 rho^2 + q5 rho^4 + q6 rho^6
//       (utilities.cpp:199-205; on the surface c = (22/5, 8/5, 2/15)/sqrt(pi),
//       which reproduces Stokes_SL_7ord_onSurf, utilities.cpp:236-252)
//   DL: s2 as P2 above, s3 = (erf + P3 rho e^{-rho^2}) / r^5,
//       P3 = q7 + q8 rho^2 + q9 rho^4 + q10 rho^6 + q11 rho^8 (utilities.cpp:280-282)
//       on the surface P3 = 2/sqrt(pi) (-1 - 2/3 rho^2 + 52/45 rho^4 - 8/45 rho^6)
//       (utilities.cpp:310-315).
constexpr int kNumCoef = 12;

// Fills the 12 polynomial coefficients from c1, c2, c3 (a target stores c,
// the sum kernels expand it when they load the target).  No contraction into
// fused multiply-adds: host and device round alike.
__host__ __device__ inline void poly_coefs(const double c[3], double q[kNumCoef]) {
#pragma clang fp contract(off)
  const double rp = 1.7724538509055160273;  // sqrt(pi)
  double c1 = c[0], c2 = c[1], c3 = c[2];
  q[0] = c1;
  q[1] = -2.0 * (c2 + c3);
  q[2] = 4.0 * c3;
  q[3] = -2.0 / rp;
  q[4] = 2.0 * (c1 + 2.0 * (c2 + c3));
  q[5] = -4.0 * (c2 + 5.0 * c3);
  q[6] = 8.0 * c3;
  q[7] = -2.0 / rp;
  q[8] = -2.0 / rp * (2.0 / 3.0);
  q[9] = 4.0 / 3.0 * (c1 + 4.0 * c2 + 12.0 * c3);
  q[10] = -8.0 / 3.0 * (c2 + 9.0 * c3);
  q[11] = 16.0 / 3.0 * c3;
}


// ---------------------------------------------------------------------------
// Device math.

// 1/sqrt(x) in FP64: the hardware estimate y (relative error below 6e-8)
// refined by one Halley step, y <- y (1 + e/2 + 3 e^2/8) with e = 1 - x y^2,
// which is the third-order Taylor step of (1 - e)^(-1/2); the residual error
// is O(e^3), far below the FP64 rounding, so the result is within about 1 ulp
// for normal x > 0.
__device__ __forceinline__ double rsqrt64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// An FP64 constant kept in vector registers for a loop.  A VALU instruction
// of this ISA reads at most one scalar register and no 64-bit literal, so an
// FMA with two non-inline constants needs one of them in vector registers;
// the empty asm keeps the compiler from rebuilding it (two moves) at every use.
__device__ __forceinline__ double vreg(double v) {
  asm("" : "+v"(v));
  return v;
}

// x^(-3/2) and x^(-5/2) directly from the hardware estimate y of x^(-1/2):
// with e = 1 - x y^2 (|e| < 1.1e-7), x^(-k/2) = y^k (1 - e)^(-k/2) and the
// binomial series to e^2, (1 + 3/2 e + 15/8 e^2) and (1 + 5/2 e + 35/8 e^2);
// the e^3 term is below 1e-20 relative, and each costs one instruction fewer
// than 1/sqrt(x) refined and raised to the power.  c2 is the e^2 coefficient
// (15/8 or 35/8) held in vector registers by the caller (vreg, outside its
// loops): with the e coefficient as the FMA's scalar addend the compiler
// emits one three-address FMA, where a vector-register addend would be tied
// to the result and copied at every use.
__device__ __forceinline__ double rsqrt3_64(double x, double c2) {
  const double y = __builtin_amdgcn_rsq(x);
  const double y2 = y * y;
  const double e = __builtin_fma(-x, y2, 1.0);
  const double y3 = y2 * y;
  return __builtin_fma(y3 * e, __builtin_fma(e, c2, 1.5), y3);
}
__device__ __forceinline__ double rsqrt5_64(double x, double c2) {
  const double y = __builtin_amdgcn_rsq(x);
  const double y2 = y * y;
  const double e = __builtin_fma(-x, y2, 1.0);
  const double y5 = y2 * y2 * y;
  return __builtin_fma(y5 * e, __builtin_fma(e, c2, 2.5), y5);
}

__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }


// LDS doubles of a sum kernel's table: the on-surface stresslet's G for
// DL_ON, erf(rho)/rho and exp(-rho^2) for every other kind.
constexpr int kDlOnTabLen = kDlOnIntervals * (kDlOnDeg + 1);
constexpr int kAETabLen = kAEIntervals * 2 * (kAEDeg + 1);
__host__ __device__ inline int table_len(int kind) { return kind == HOR_DL_ON ? kDlOnTabLen : kAETabLen; }
// In LDS the tables are coefficient-major: coefficient j of every interval is
// contiguous, so the lanes of a wave, whose arguments fall in neighbouring
// intervals, read neighbouring words (distinct LDS banks, or one broadcast).
constexpr int kAERow = 2 * (kAEDeg + 1);   // doubles per interval in erf_exp_table.h (A, then E)

// Source record read by the sum kernels, one per sorted source: position,
// weight a, a-scaled normal a n, and the density, a-scaled for the single layer
// (a f) and plain for the double layer (g).  Every pair term is linear in the
// normal (DL) or in the density (SL), so the weight folds in exactly.
struct __align__(16) SrcRec {
  double x, y, z, a;
  double nx, ny, nz, pad0;
  double d0, d1, d2, pad1;
};

// Record of the fused Stokeslet + stresslet sum: position, a n, a f, g.
struct __align__(16) SrcRecV {
  double x, y, z, nx;
  double ny, nz, f0, f1;
  double f2, g0, g1, g2;
};

// FP32 copy of SrcRecV (mixed precision), position relative to the centre of
// the source's 64-subtile box.
struct __align__(16) SrcRecV32 {
  float x, y, z, nx;
  float ny, nz, f0, f1;
  float f2, g0, g1, g2;
};

// Record of the regularized pressure (HOR_P_OFF): position x' = x - c relative
// to the centre c of the source's 64-subtile box, a n, a f, al = a n.x',
// be = a f.x', m = a n x x' and xx = |x'|^2, which turn the plain pair's r^2
// and numerator into chains of fused multiply-adds (p_plain).
struct __align__(16) SrcRecP {
  double x, y, z, al;
  double nx, ny, nz, be;
  double d0, d1, d2, m0;
  double m1, m2, xx, pad;
};

struct __align__(16) SrcRec32 {
  float x, y, z, a;
  float nx, ny, nz, pad0;
  float d0, d1, d2, pad1;
};

}  // namespace hor
