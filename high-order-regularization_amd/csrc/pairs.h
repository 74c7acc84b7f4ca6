// Pair functions shared by the direct-sum and treecode leaf kernels.
//
// Pair arithmetic follows the reference's per-pair expressions
// (highreg_stokes_2surf/utilities.cpp:179-323, stokeslets.cpp:106-130,
// stresslets.cpp:117-168, 244-269).
#pragma once

#include "device.h"
#include "direct.h"

namespace hor {

// Stages the kind's table in LDS; every thread of the workgroup must call it
// before any of them returns.
template <int KIND>
__device__ __forceinline__ const double* stage_table(const DirectArgs& A) {
  extern __shared__ double lds_tab[];
  const double* src = KIND == HOR_DL_ON ? A.dlon_tab : A.ae_tab;
  if (KIND == HOR_DL_ON) {
    for (int q = threadIdx.x; q < kDlOnTabLen; q += blockDim.x)
      lds_tab[(q % (kDlOnDeg + 1)) * kDlOnIntervals + q / (kDlOnDeg + 1)] = src[q];
  } else {
    for (int q = threadIdx.x; q < kAETabLen; q += blockDim.x) lds_tab[(q % kAERow) * kAEIntervals + q / kAERow] = src[q];
  }
  __syncthreads();
  return lds_tab;
}

// A = erf(rho)/rho and E = exp(-rho^2) at s = rho^2 <= 64 from the LDS table
// (erf_exp_table.h): interval k = floor(s / w), local x in [-1, 1].
__device__ __forceinline__ void erf_exp_AE(const double* __restrict__ tab, double s, double& A, double& E) {
  const int k = min((int)(s * (kAEIntervals / kAESmax)), kAEIntervals - 1);
  const double x = s * (2.0 * kAEIntervals / kAESmax) - (double)(2 * k + 1);
  const double* c = tab + k;
  constexpr int S = kAEIntervals;
  double a = c[kAEDeg * S], e = c[(2 * kAEDeg + 1) * S];
#pragma unroll
  for (int j = kAEDeg - 1; j >= 0; j--) {
    a = __builtin_fma(a, x, c[j * S]);
    e = __builtin_fma(e, x, c[(kAEDeg + 1 + j) * S]);
  }
  A = a;
  E = e;
}

// G(s) = F(rho)/rho^5 of the on-surface kernel from its piecewise polynomial
// (dl_on_table.h), tab in LDS, at v = s / w (the interval coordinate, w the
// interval width): interval k = floor(v), local x = 2 v - (2 k + 1) in [-1, 1].
__device__ __forceinline__ double dl_on_G(const double* __restrict__ tab, double v) {
  const int k = min((int)v, kDlOnIntervals - 1);
  const double x = 2.0 * v - (double)(2 * k + 1);
  const double* c = tab + k;
  constexpr int S = kDlOnIntervals;
  double g = c[kDlOnDeg * S];
#pragma unroll
  for (int j = kDlOnDeg - 1; j >= 0; j--) g = __builtin_fma(g, x, c[j * S]);
  return g;
}


// The active lanes of the wave where the predicate holds (a scalar mask).  The
// per-source votes take the ballot of the distance compare alone and AND it
// with the loop-invariant mask of near targets, so a vote is one vector
// compare and scalar instructions.
__device__ __forceinline__ unsigned long long wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Box distance test between a wave's target box and a source subtile box.
__device__ __forceinline__ bool boxes_far(const double* wb, const double* sb, double far_d2) {
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double g = fmax(fmax(wb[k] - sb[3 + k], sb[k] - wb[3 + k]), 0.0);
    d2 += g * g;
  }
  return d2 > far_d2;
}

struct Tgt {
  double px, py, pz;
  double sub;           // SL: f0.n0 (0 for a far target)
  double g0x, g0y, g0z; // DL: g0 (0 for a far target)
  double x0x, x0y, x0z, n0x, n0y, n0z, b;  // DL_OFF
  double q[kNumCoef];
  bool nearp;
};

// Source records carry a n and, for the single layer, a f (SrcRec), so no
// pair term multiplies by the weight a.

// ---- SL pair: plain kernel with subtraction a [f - (f0.n0) n] (fast path):
//   u += a f~ / r + dx (a f~ . dx) / r^3  =  (a f~ + dx t) / r,  t = (a f~ . dx) / r^2
// with f~ . dx = f.dx - (f0.n0)(n.dx) and a f~ + dx t = (a f + dx t) - (f0.n0) a n:
// every FMA then reads one source value, which the wave holds in scalar
// registers, and f~ is never formed (a VALU instruction reads at most one
// scalar register, so f - (f0.n0) n would first copy f to vector registers).
__device__ __forceinline__ void sl_plain(const SrcRec& s, double px, double py, double pz, double sub,
                                         double u[3]) {
  double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double ri = rsqrt64(r2);
  double fdx = s.d0 * dx + s.d1 * dy + s.d2 * dz;
  double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
  double t = (fdx - sub * ndx) * (ri * ri);
  u[0] += ((s.d0 + dx * t) - sub * s.nx) * ri;
  u[1] += ((s.d1 + dy * t) - sub * s.ny) * ri;
  u[2] += ((s.d2 + dz * t) - sub * s.nz) * ri;
}

// The same in FP32 for a given separation d (FP32 partial sums).
__device__ __forceinline__ void sl_plain32(const SrcRec32& s, float dx, float dy, float dz, float sub,
                                           float u[3]) {
  float r2 = dx * dx + dy * dy + dz * dz;
  float ri = rsqrt32(r2);
  float fx = s.d0 - sub * s.nx, fy = s.d1 - sub * s.ny, fz = s.d2 - sub * s.nz;
  float t = (fx * dx + fy * dy + fz * dz) * (ri * ri);
  u[0] += (fx + dx * t) * ri;
  u[1] += (fy + dy * t) * ri;
  u[2] += (fz + dz * t) * ri;
}

// ---- SL pair, full branch structure (mixed path), FP64
// (utilities.cpp:179-207, stokeslets.cpp:106-130).  r = sqrt(r2) and 1/r come
// from one reciprocal square root; the self test r < 1e-14 is r2 < 1e-28.
__device__ __forceinline__ void sl_full(const SrcRec& s, double px, double py, double pz,
                                        double sub, bool nearp, const double* q, double inv_del, double inv_del2,
                                        double self1, double self2, const double* __restrict__ tab, double u[3]) {
  double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double fx = s.d0 - sub * s.nx, fy = s.d1 - sub * s.ny, fz = s.d2 - sub * s.nz;
  double fdx = fx * dx + fy * dy + fz * dz;
  double s1, s2;
  if (nearp && r2 < 1e-28) {
    s1 = self1;
    s2 = self2;
  } else {
    const double ri = rsqrt64(r2);
    const double sq = r2 * inv_del2;   // rho^2
    if (!nearp || sq > 64.0) {
      s1 = ri;
      s2 = ri * ri * ri;
    } else {
      // erf(rho) = rho A, rho e^{-rho^2} = rho E and rho / r = 1/delta
      double A, E;
      erf_exp_AE(tab, sq, A, E);
      const double p1 = q[0] + sq * (q[1] + sq * q[2]);
      const double p2 = q[3] + sq * (q[4] + sq * (q[5] + sq * q[6]));
      s1 = (A + p1 * E) * inv_del;
      s2 = (A + p2 * E) * inv_del * (ri * ri);
    }
  }
  double H2 = s2 * fdx;
  u[0] += fx * s1 + dx * H2;
  u[1] += fy * s1 + dy * H2;
  u[2] += fz * s1 + dz * H2;
}

// ---- DL pair: plain kernel (g - g0).d (a n.d) d / r^5 (fast path).  For a
// target within 8 delta of the surface this is the reference's
// subtraction form (stresslets.cpp:157-167) evaluated in the algebraically
// identical, cancellation-free product form.
__device__ __forceinline__ void dl_plain(const SrcRec& s, double px, double py, double pz, double g0x,
                                         double g0y, double g0z, double k5, double u[3]) {
  double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double gdx = (s.d0 - g0x) * dx + (s.d1 - g0y) * dy + (s.d2 - g0z) * dz;
  double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
  double k = gdx * ndx * rsqrt5_64(r2, k5);
  u[0] += dx * k;
  u[1] += dy * k;
  u[2] += dz * k;
}

// The same in FP32 for a given separation d (FP32 partial sums).
__device__ __forceinline__ void dl_plain32(const SrcRec32& s, float dx, float dy, float dz, float g0x, float g0y,
                                           float g0z, float u[3]) {
  float r2 = dx * dx + dy * dy + dz * dz;
  float ri = rsqrt32(r2);
  float ri2 = ri * ri;
  float gdx = (s.d0 - g0x) * dx + (s.d1 - g0y) * dy + (s.d2 - g0z) * dz;
  float ndx = s.nx * dx + s.ny * dy + s.nz * dz;
  float k = gdx * ndx * (ri2 * ri2 * ri);
  u[0] += dx * k;
  u[1] += dy * k;
  u[2] += dz * k;
}

// ---- DL off-surface pair, full branch structure (stresslets.cpp:117-168).
// Every term is linear in the source normal, so the record's a n carries the
// weight.
__device__ __forceinline__ void dl_off_full(const SrcRec& s, const Tgt& t, double inv_del, double inv_del2,
                                            double self2, const double* __restrict__ tab, double u[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double gx = s.d0 - t.g0x, gy = s.d1 - t.g0y, gz = s.d2 - t.g0z;
  double H2 = 0.0, H3;
  bool plain = false;
  if (t.nearp && r2 < 1e-28) {
    H2 = self2;
    H3 = 0.0;
  } else {
    double ri = rsqrt64(r2);
    double ri2 = ri * ri;
    const double sq = r2 * inv_del2;   // rho^2
    if (!t.nearp || sq > 64.0) {
      plain = true;
      H3 = ri2 * ri2 * ri;
    } else {
      // erf(rho) = rho A, rho e^{-rho^2} = rho E and rho / r = 1/delta
      double A, E;
      erf_exp_AE(tab, sq, A, E);
      const double* q = t.q;
      double p2 = q[3] + sq * (q[4] + sq * (q[5] + sq * q[6]));
      double p3 = q[7] + sq * (q[8] + sq * (q[9] + sq * (q[10] + sq * q[11])));
      const double w2 = inv_del * ri2;
      H2 = (A + p2 * E) * w2;
      H3 = (A + p3 * E) * (w2 * ri2);
    }
  }
  double c0, c1, c2;
  if (plain) {
    double gdx = gx * dx + gy * dy + gz * dz;
    double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
    double k = gdx * ndx * H3;
    c0 = dx * k;
    c1 = dy * k;
    c2 = dz * k;
  } else {
    double xh0 = s.x - t.x0x, xh1 = s.y - t.x0y, xh2 = s.z - t.x0z;
    double sig2 = (xh0 * xh0 + xh1 * xh1 + xh2 * xh2) - 2.0 * t.b * (xh0 * t.n0x + xh1 * t.n0y + xh2 * t.n0z);
    double g_n0 = gx * t.n0x + gy * t.n0y + gz * t.n0z;
    double g_xh = gx * xh0 + gy * xh1 + gz * xh2;
    double n0_n = t.n0x * s.nx + t.n0y * s.ny + t.n0z * s.nz;
    double xh_n = xh0 * s.nx + xh1 * s.ny + xh2 * s.nz;
    double m1a = g_n0 * n0_n;
    double m1b = g_xh * n0_n + g_n0 * xh_n;
    double m2a = g_xh * xh_n;
    double m2b = t.b * m1b;
    // t1 = d m1a - n0 m1b;  u += t1 (H2 - sig2 H3) + (d m2a + xh m2b) H3,
    // collected per vector: u += d Bd - n0 Bn + xh Bx
    double A1 = H2 - sig2 * H3;
    double Bd = m1a * A1 + m2a * H3, Bn = m1b * A1, Bx = m2b * H3;
    c0 = dx * Bd - t.n0x * Bn + xh0 * Bx;
    c1 = dy * Bd - t.n0y * Bn + xh1 * Bx;
    c2 = dz * Bd - t.n0z * Bn + xh2 * Bx;
  }
  u[0] += c0;
  u[1] += c1;
  u[2] += c2;
}

// ---- Stokeslet + stresslet at one off-surface target (HOR_SLDL_OFF): the
// pair's separation, r^2 and 1/r are computed once; the single-layer terms
// are sl_plain's / sl_full's and the double-layer terms dl_plain's /
// dl_off_full's, expression for expression; n.dx serves both.
__device__ __forceinline__ void sldl_plain(const SrcRecV& s, const Tgt& t, double us[3], double ud[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double ri = rsqrt64(r2);
  double fdx = s.f0 * dx + s.f1 * dy + s.f2 * dz;
  double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
  double tq = (fdx - t.sub * ndx) * (ri * ri);
  us[0] += ((s.f0 + dx * tq) - t.sub * s.nx) * ri;
  us[1] += ((s.f1 + dy * tq) - t.sub * s.ny) * ri;
  us[2] += ((s.f2 + dz * tq) - t.sub * s.nz) * ri;
  double ri2 = ri * ri;
  double gdx = (s.g0 - t.g0x) * dx + (s.g1 - t.g0y) * dy + (s.g2 - t.g0z) * dz;
  double k = gdx * ndx * (ri2 * ri2 * ri);
  ud[0] += dx * k;
  ud[1] += dy * k;
  ud[2] += dz * k;
}

__device__ __forceinline__ void sldl_full(const SrcRecV& s, const Tgt& t, double inv_del, double inv_del2,
                                          double self1, double self2, const double* __restrict__ tab, double us[3],
                                          double ud[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double fx = s.f0 - t.sub * s.nx, fy = s.f1 - t.sub * s.ny, fz = s.f2 - t.sub * s.nz;
  double fdx = fx * dx + fy * dy + fz * dz;
  double gx = s.g0 - t.g0x, gy = s.g1 - t.g0y, gz = s.g2 - t.g0z;
  double s1, s2, H2 = 0.0, H3;
  bool plain = false;
  if (t.nearp && r2 < 1e-28) {
    s1 = self1;
    s2 = self2;
    H2 = self2;
    H3 = 0.0;
  } else {
    const double ri = rsqrt64(r2);
    const double ri2 = ri * ri;
    const double sq = r2 * inv_del2;   // rho^2
    if (!t.nearp || sq > 64.0) {
      s1 = ri;
      s2 = ri * ri * ri;
      plain = true;
      H3 = ri2 * ri2 * ri;
    } else {
      double A, E;
      erf_exp_AE(tab, sq, A, E);
      const double* q = t.q;
      const double p1 = q[0] + sq * (q[1] + sq * q[2]);
      const double p2 = q[3] + sq * (q[4] + sq * (q[5] + sq * q[6]));
      const double p3 = q[7] + sq * (q[8] + sq * (q[9] + sq * (q[10] + sq * q[11])));
      s1 = (A + p1 * E) * inv_del;
      // the Stokeslet's s2 and the stresslet's H2 are one function (Stokes_SL_7ord
      // and Stokes_DL_7ord, utilities.cpp:199-205, 276-285)
      const double w2 = inv_del * ri2;
      s2 = (A + p2 * E) * w2;
      H2 = s2;
      H3 = (A + p3 * E) * (w2 * ri2);
    }
  }
  double HS = s2 * fdx;
  us[0] += fx * s1 + dx * HS;
  us[1] += fy * s1 + dy * HS;
  us[2] += fz * s1 + dz * HS;
  double c0, c1, c2;
  if (plain) {
    double gdx = gx * dx + gy * dy + gz * dz;
    double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
    double k = gdx * ndx * H3;
    c0 = dx * k;
    c1 = dy * k;
    c2 = dz * k;
  } else {
    double xh0 = s.x - t.x0x, xh1 = s.y - t.x0y, xh2 = s.z - t.x0z;
    double sig2 = (xh0 * xh0 + xh1 * xh1 + xh2 * xh2) - 2.0 * t.b * (xh0 * t.n0x + xh1 * t.n0y + xh2 * t.n0z);
    double g_n0 = gx * t.n0x + gy * t.n0y + gz * t.n0z;
    double g_xh = gx * xh0 + gy * xh1 + gz * xh2;
    double n0_n = t.n0x * s.nx + t.n0y * s.ny + t.n0z * s.nz;
    double xh_n = xh0 * s.nx + xh1 * s.ny + xh2 * s.nz;
    double m1a = g_n0 * n0_n;
    double m1b = g_xh * n0_n + g_n0 * xh_n;
    double m2a = g_xh * xh_n;
    double m2b = t.b * m1b;
    double A1 = H2 - sig2 * H3;
    double Bd = m1a * A1 + m2a * H3, Bn = m1b * A1, Bx = m2b * H3;
    c0 = dx * Bd - t.n0x * Bn + xh0 * Bx;
    c1 = dy * Bd - t.n0y * Bn + xh1 * Bx;
    c2 = dz * Bd - t.n0z * Bn + xh2 * Bx;
  }
  ud[0] += c0;
  ud[1] += c1;
  ud[2] += c2;
}

// ---- DL on-surface pair, full branch structure (stresslets.cpp:244-269).
// On-surface stresslet pair (Stokes_DL_7ord_onSurf, utilities.cpp:291-323):
// H3 = F(rho)/r^5 with F = erf(rho) + p3(rho^2) rho e^{-rho^2}; within 8 delta
// H3 = G(rho^2)/delta^5 from the table (no exp, erf or square root), beyond
// it the plain 1/r^5.
// vsc = (1/delta^2) / w, w = kDlOnSmax / kDlOnIntervals (a power of two), so
// v = r^2 vsc is (r/delta)^2 / w exactly and the 8 delta test is v > 64 / w.
__device__ __forceinline__ void dl_on_full(const SrcRec& s, const Tgt& t, double vsc, double inv_del5,
                                           const double* __restrict__ tab, double k5, double u[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 < 1e-28) return;  // d = 0: the self term contributes exactly zero
  const double v = r2 * vsc;
  double H3;
  if (v > 64.0 * (kDlOnIntervals / kDlOnSmax)) {
    H3 = rsqrt5_64(r2, k5);
  } else {
    H3 = dl_on_G(tab, v) * inv_del5;
  }
  double gx = s.d0 - t.g0x, gy = s.d1 - t.g0y, gz = s.d2 - t.g0z;
  double dxn = dx * s.nx + dy * s.ny + dz * s.nz;
  double dxg = dx * gx + dy * gy + dz * gz;
  H3 *= dxn * dxg;
  u[0] += dx * H3;
  u[1] += dy * H3;
  u[2] += dz * H3;
}

// ---------------------------------------------------------------------------
// Mixed precision: regularized pairs with 4 delta < r <= 8 delta in FP32.
// Every lane of the wave is beyond 4 delta of the source when these run, so
// there is no self term, erf(rho) is within 2e-8 of 1 and the correction
// terms P(rho^2) rho exp(-rho^2) are below 3e-2 of the kernel.

// Per-target data of the FP32 pairs (converted once per target).
struct Tgt32 {
  float sub;                 // SL: f0.n0
  float g0x, g0y, g0z;       // DL: g0
  float ex, ey, ez;          // DL_OFF: p - x0, so that x^ = y - x0 = e - d
  float n0x, n0y, n0z, b;    // DL_OFF
  float q[kNumCoef];
};

__device__ __forceinline__ void make_tgt32(const Tgt& t, Tgt32& v) {
  v.sub = (float)t.sub;
  v.g0x = (float)t.g0x; v.g0y = (float)t.g0y; v.g0z = (float)t.g0z;
  v.ex = (float)(t.px - t.x0x); v.ey = (float)(t.py - t.x0y); v.ez = (float)(t.pz - t.x0z);
  v.n0x = (float)t.n0x; v.n0y = (float)t.n0y; v.n0z = (float)t.n0z; v.b = (float)t.b;
#pragma unroll
  for (int k = 0; k < kNumCoef; k++) v.q[k] = (float)t.q[k];
}

// erf(rho) for 4 < rho <= 8 from E = exp(-rho^2): 1 - E erfcx(rho), with the
// asymptotic erfcx(rho) ~ (1 - 1/(2 rho^2) + 3/(4 rho^4)) / (sqrt(pi) rho),
// whose relative error (below 5e-4 at rho = 4) moves erf by less than 1e-11.
__device__ __forceinline__ float erf_tail32(float rho, float E) {
  const float ir = __builtin_amdgcn_rcpf(rho);
  const float ir2 = ir * ir;
  return 1.0f - E * (0.5641895835f * ir) * (1.0f + ir2 * (-0.5f + 0.75f * ir2));
}

// exp(-x) in FP32 for 0 <= x <= 64 (v_exp_f32 of -x log2 e).
__device__ __forceinline__ float expneg32(float x) { return __builtin_amdgcn_exp2f(-1.4426950408889634f * x); }

__device__ __forceinline__ void sl_reg32(const SrcRec32& s, float dx, float dy, float dz, const Tgt32& v,
                                         bool nearp, float inv_del, float u[3]) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = rsqrt32(r2);
  const float rd = (r2 * ri) * inv_del;
  const float ri3 = ri * ri * ri;
  const float fx = s.d0 - v.sub * s.nx, fy = s.d1 - v.sub * s.ny, fz = s.d2 - v.sub * s.nz;
  const float fdx = fx * dx + fy * dy + fz * dz;
  float s1 = ri, s2 = ri3;
  if (nearp && !(rd > 8.0f)) {
    const float rd2 = rd * rd;
    const float E = expneg32(rd2);
    const float ef = erf_tail32(rd, E);
    const float rex = rd * E;
    const float p1 = v.q[0] + rd2 * (v.q[1] + rd2 * v.q[2]);
    const float p2 = v.q[3] + rd2 * (v.q[4] + rd2 * (v.q[5] + rd2 * v.q[6]));
    s1 = (ef + p1 * rex) * ri;
    s2 = (ef + p2 * rex) * ri3;
  }
  const float H2 = s2 * fdx;
  u[0] += fx * s1 + dx * H2;
  u[1] += fy * s1 + dy * H2;
  u[2] += fz * s1 + dz * H2;
}

__device__ __forceinline__ void dl_off_reg32(const SrcRec32& s, float dx, float dy, float dz, const Tgt32& v,
                                             bool nearp, float inv_del, float u[3]) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = rsqrt32(r2);
  const float rd = (r2 * ri) * inv_del;
  const float ri2 = ri * ri;
  const float gx = s.d0 - v.g0x, gy = s.d1 - v.g0y, gz = s.d2 - v.g0z;
  if (!nearp || rd > 8.0f) {
    const float gdx = gx * dx + gy * dy + gz * dz;
    const float ndx = s.nx * dx + s.ny * dy + s.nz * dz;
    const float k = gdx * ndx * (ri2 * ri2 * ri);
    u[0] += dx * k;
    u[1] += dy * k;
    u[2] += dz * k;
    return;
  }
  const float ri3 = ri2 * ri;
  const float rd2 = rd * rd;
  const float E = expneg32(rd2);
  const float ef = erf_tail32(rd, E);
  const float rex = rd * E;
  const float p2 = v.q[3] + rd2 * (v.q[4] + rd2 * (v.q[5] + rd2 * v.q[6]));
  const float p3 = v.q[7] + rd2 * (v.q[8] + rd2 * (v.q[9] + rd2 * (v.q[10] + rd2 * v.q[11])));
  const float H2 = (ef + p2 * rex) * ri3;
  const float H3 = (ef + p3 * rex) * (ri3 * ri2);
  const float xh0 = v.ex - dx, xh1 = v.ey - dy, xh2 = v.ez - dz;
  const float sig2 = (xh0 * xh0 + xh1 * xh1 + xh2 * xh2) - 2.0f * v.b * (xh0 * v.n0x + xh1 * v.n0y + xh2 * v.n0z);
  const float g_n0 = gx * v.n0x + gy * v.n0y + gz * v.n0z;
  const float g_xh = gx * xh0 + gy * xh1 + gz * xh2;
  const float n0_n = v.n0x * s.nx + v.n0y * s.ny + v.n0z * s.nz;
  const float xh_n = xh0 * s.nx + xh1 * s.ny + xh2 * s.nz;
  const float m1a = g_n0 * n0_n;
  const float m1b = g_xh * n0_n + g_n0 * xh_n;
  const float m2a = g_xh * xh_n;
  const float m2b = v.b * m1b;
  const float A1 = H2 - sig2 * H3;
  const float Bd = m1a * A1 + m2a * H3, Bn = m1b * A1, Bx = m2b * H3;
  u[0] += dx * Bd - v.n0x * Bn + xh0 * Bx;
  u[1] += dy * Bd - v.n0y * Bn + xh1 * Bx;
  u[2] += dz * Bd - v.n0z * Bn + xh2 * Bx;
}

// Stokeslet + stresslet (HOR_SLDL_OFF) in FP32: sl_plain32 / dl_plain32 and
// sl_reg32 / dl_off_reg32 with the separation, 1/r, rho, exp(-rho^2), erf(rho)
// and the s2 polynomial computed once.
__device__ __forceinline__ void sldl_plain32(const SrcRecV32& s, float dx, float dy, float dz, const Tgt32& v,
                                             float ws[3], float wd[3]) {
  float r2 = dx * dx + dy * dy + dz * dz;
  float ri = rsqrt32(r2);
  float fdx = s.f0 * dx + s.f1 * dy + s.f2 * dz;
  float ndx = s.nx * dx + s.ny * dy + s.nz * dz;
  float tq = (fdx - v.sub * ndx) * (ri * ri);
  ws[0] += ((s.f0 + dx * tq) - v.sub * s.nx) * ri;
  ws[1] += ((s.f1 + dy * tq) - v.sub * s.ny) * ri;
  ws[2] += ((s.f2 + dz * tq) - v.sub * s.nz) * ri;
  float ri2 = ri * ri;
  float gdx = (s.g0 - v.g0x) * dx + (s.g1 - v.g0y) * dy + (s.g2 - v.g0z) * dz;
  float k = gdx * ndx * (ri2 * ri2 * ri);
  wd[0] += dx * k;
  wd[1] += dy * k;
  wd[2] += dz * k;
}

__device__ __forceinline__ void sldl_reg32(const SrcRecV32& s, float dx, float dy, float dz, const Tgt32& v,
                                           bool nearp, float inv_del, float ws[3], float wd[3]) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = rsqrt32(r2);
  const float rd = (r2 * ri) * inv_del;
  const float ri2 = ri * ri;
  const float ri3 = ri * ri * ri;
  const float fx = s.f0 - v.sub * s.nx, fy = s.f1 - v.sub * s.ny, fz = s.f2 - v.sub * s.nz;
  const float fdx = fx * dx + fy * dy + fz * dz;
  const float gx = s.g0 - v.g0x, gy = s.g1 - v.g0y, gz = s.g2 - v.g0z;
  if (!nearp || rd > 8.0f) {
    const float HS = ri3 * fdx;
    ws[0] += fx * ri + dx * HS;
    ws[1] += fy * ri + dy * HS;
    ws[2] += fz * ri + dz * HS;
    const float gdx = gx * dx + gy * dy + gz * dz;
    const float ndx = s.nx * dx + s.ny * dy + s.nz * dz;
    const float k = gdx * ndx * (ri2 * ri2 * ri);
    wd[0] += dx * k;
    wd[1] += dy * k;
    wd[2] += dz * k;
    return;
  }
  const float rd2 = rd * rd;
  const float E = expneg32(rd2);
  const float ef = erf_tail32(rd, E);
  const float rex = rd * E;
  const float p1 = v.q[0] + rd2 * (v.q[1] + rd2 * v.q[2]);
  const float p2 = v.q[3] + rd2 * (v.q[4] + rd2 * (v.q[5] + rd2 * v.q[6]));
  const float p3 = v.q[7] + rd2 * (v.q[8] + rd2 * (v.q[9] + rd2 * (v.q[10] + rd2 * v.q[11])));
  const float s1 = (ef + p1 * rex) * ri;
  const float s2 = (ef + p2 * rex) * ri3;
  const float HS = s2 * fdx;
  ws[0] += fx * s1 + dx * HS;
  ws[1] += fy * s1 + dy * HS;
  ws[2] += fz * s1 + dz * HS;
  const float H2 = (ef + p2 * rex) * (ri2 * ri);
  const float H3 = (ef + p3 * rex) * ((ri2 * ri) * ri2);
  const float xh0 = v.ex - dx, xh1 = v.ey - dy, xh2 = v.ez - dz;
  const float sig2 = (xh0 * xh0 + xh1 * xh1 + xh2 * xh2) - 2.0f * v.b * (xh0 * v.n0x + xh1 * v.n0y + xh2 * v.n0z);
  const float g_n0 = gx * v.n0x + gy * v.n0y + gz * v.n0z;
  const float g_xh = gx * xh0 + gy * xh1 + gz * xh2;
  const float n0_n = v.n0x * s.nx + v.n0y * s.ny + v.n0z * s.nz;
  const float xh_n = xh0 * s.nx + xh1 * s.ny + xh2 * s.nz;
  const float m1a = g_n0 * n0_n;
  const float m1b = g_xh * n0_n + g_n0 * xh_n;
  const float m2a = g_xh * xh_n;
  const float m2b = v.b * m1b;
  const float A1 = H2 - sig2 * H3;
  const float Bd = m1a * A1 + m2a * H3, Bn = m1b * A1, Bx = m2b * H3;
  wd[0] += dx * Bd - v.n0x * Bn + xh0 * Bx;
  wd[1] += dy * Bd - v.n0y * Bn + xh1 * Bx;
  wd[2] += dz * Bd - v.n0z * Bn + xh2 * Bx;
}

__device__ __forceinline__ void dl_on_reg32(const SrcRec32& s, float dx, float dy, float dz, const Tgt32& v,
                                            float inv_del, float u[3]) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = rsqrt32(r2);
  const float rd = (r2 * ri) * inv_del;
  const float ri2 = ri * ri;
  float H3 = ri2 * ri2 * ri;
  if (!(rd > 8.0f)) {
    const float rd2 = rd * rd;
    const float E = expneg32(rd2);
    const float ef = erf_tail32(rd, E);
    const float p3 = v.q[7] + rd2 * (v.q[8] + rd2 * (v.q[9] + rd2 * v.q[10]));
    H3 = (ef + p3 * (rd * E)) * H3;
  }
  const float gx = s.d0 - v.g0x, gy = s.d1 - v.g0y, gz = s.d2 - v.g0z;
  const float k = H3 * (dx * s.nx + dy * s.ny + dz * s.nz) * (dx * gx + dy * gy + dz * gz);
  u[0] += dx * k;
  u[1] += dy * k;
  u[2] += dz * k;
}

// ---- Pressure pair (stokes_pressure_7ord, extend_to_grid/stokes.cpp:134-203).
// The reference's summand [-(f.n - f0.n0)(n.dx) - (n x f - n0 x f0).(n x dx)] A
// equals dx.w with w = -(a f) + (f0.n0)(a n) + (n0 x f0) x (a n) for a unit
// normal n (Binet-Cauchy and the triple product), i.e.
//   (f0.n0)(a n . dx) - a f . dx + a n . (dx x (n0 x f0)).
// Tgt.sub holds f0.n0 and Tgt.g0 holds n0 x f0.  Every pair evaluates it in
// its source subtile's frame as one FMA chain (p_chain, below): each product
// reads at most one source value (the wave holds the record in scalar
// registers and a VALU instruction reads at most one).

// Per-target data of the pressure's plain pairs in the frame of one source
// subtile (box centre c): p' = p - c, m2 = -2 p', pp = |p'|^2 and
// q = (f0.n0) p' + p' x (n0 x f0).
struct TgtP {
  double px, py, pz, qx, qy, qz, m2x, m2y, m2z, pp;
};

__device__ __forceinline__ void make_tgtp(const Tgt& t, const double* box, TgtP& v) {
  v.px = t.px - 0.5 * (box[0] + box[3]);
  v.py = t.py - 0.5 * (box[1] + box[4]);
  v.pz = t.pz - 0.5 * (box[2] + box[5]);
  v.qx = t.sub * v.px + (v.py * t.g0z - v.pz * t.g0y);
  v.qy = t.sub * v.py + (v.pz * t.g0x - v.px * t.g0z);
  v.qz = t.sub * v.pz + (v.px * t.g0y - v.py * t.g0x);
  v.m2x = -2.0 * v.px;
  v.m2y = -2.0 * v.py;
  v.m2z = -2.0 * v.pz;
  v.pp = v.px * v.px + v.py * v.py + v.pz * v.pz;
}

// The numerator in the subtile's frame (d = p' - x'): (f0.n0)(n.d) - f.d +
// n.(d x c), c = n0 x f0, is n.q - f.p' - c.m + (be - (f0.n0) al), one FMA
// chain on the record's precomputed al, be, m.  For the plain kernel H2 = 1/r^3
// (Lapl_DL_7ord beyond 8 delta) r^2 = (|p'|^2 + |x'|^2) - 2 p'.x' is a second
// chain on the record's xx.  The terms of both are of size |p'| + |x'|, where
// |x'| is within the subtile's box and |p'| - |x'| <= r, so their rounding
// stays within a few ulps times (|p'| + |x'|) / r: a few ulps for the plain
// pairs (r > 8 delta), a few 1e-15 relative for the closest regularized ones.
__device__ __forceinline__ double p_chain(const SrcRecP& s, const Tgt& t, const TgtP& v) {
  double a = v.qx * s.nx;
  a = __builtin_fma(v.qy, s.ny, a);
  a = __builtin_fma(v.qz, s.nz, a);
  a = __builtin_fma(-v.px, s.d0, a);
  a = __builtin_fma(-v.py, s.d1, a);
  a = __builtin_fma(-v.pz, s.d2, a);
  a = __builtin_fma(-t.g0x, s.m0, a);
  a = __builtin_fma(-t.g0y, s.m1, a);
  a = __builtin_fma(-t.g0z, s.m2, a);
  a = __builtin_fma(-t.sub, s.al, a);
  return a + s.be;
}

__device__ __forceinline__ void p_plain_r2(const SrcRecP& s, const Tgt& t, const TgtP& v, double r2, double k3,
                                           double u[3]) {
  u[0] += p_chain(s, t, v) * rsqrt3_64(r2, k3);
}

__device__ __forceinline__ void p_plain(const SrcRecP& s, const Tgt& t, const TgtP& v, double k3, double u[3]) {
  double r2 = v.pp + s.xx;
  r2 = __builtin_fma(v.m2x, s.x, r2);
  r2 = __builtin_fma(v.m2y, s.y, r2);
  r2 = __builtin_fma(v.m2z, s.z, r2);
  p_plain_r2(s, t, v, r2, k3, u);
}

// Full branch structure of Lapl_DL_7ord (extend_to_grid/utilities.cpp:190-211)
// at a pair of a mixed subtile, r^2 from the vote: plain, or
// (erf + P2(rho^2) rho e^{-rho^2}) / r^3 with the Stokes_SL_7ord s2 polynomial
// (the same coefficients q[3..6]), times the numerator as the plain pair's FMA
// chain in the subtile frame.  Its terms are of size |p'| + |x'| <= r + 2 x
// (the box radius), so a near pair's numerator is within a few ulps times
// (|p'| + |x'|) / r of d.w (a few 1e-15 at the closest band pairs).
__device__ __forceinline__ void p_full(const SrcRecP& s, const Tgt& t, const TgtP& v, double r2, double inv_del,
                                       double inv_del2, const double* __restrict__ tab, double u[3]) {
  // d = 0: the numerator, linear in d, vanishes (the reference multiplies it
  // by the self limit of H2)
  if (t.nearp && r2 < 1e-28) return;
  double H2;
  {
    const double ri = rsqrt64(r2);
    const double sq = r2 * inv_del2;   // rho^2
    if (!t.nearp || sq > 64.0) {
      H2 = ri * ri * ri;
    } else {
      double A, E;
      erf_exp_AE(tab, sq, A, E);
      const double p2 = t.q[3] + sq * (t.q[4] + sq * (t.q[5] + sq * t.q[6]));
      H2 = (A + p2 * E) * (inv_del * (ri * ri));
    }
  }
  u[0] += p_chain(s, t, v) * H2;
}

// Target data of sum kernel KIND.  NOSUB (double layer): the pair terms run on
// g instead of g - g0, the subtraction being applied afterwards through the
// plan's precomputed coefficients (tree plans, C_i g0_i in k_tree_finalize).
template <int KIND, bool NOSUB = false>
__device__ __forceinline__ void load_target(const DirectArgs& A, int i, Tgt& t) {
  int ntp = A.ntp;
  t.px = A.tx[i];
  t.py = A.tx[ntp + i];
  t.pz = A.tx[2 * ntp + i];
  t.nearp = A.tnear[i] != 0;
  if (KIND != HOR_DL_ON && t.nearp) {
    const double c[3] = {A.tcoef[i], A.tcoef[ntp + i], A.tcoef[2 * ntp + i]};
    poly_coefs(c, t.q);
  } else {
#pragma unroll
    for (int k = 0; k < kNumCoef; k++) t.q[k] = 0.0;
  }
  t.sub = 0;
  t.g0x = t.g0y = t.g0z = 0;
  t.x0x = t.x0y = t.x0z = t.n0x = t.n0y = t.n0z = t.b = 0;
  if (KIND == HOR_SL_ON || KIND == HOR_SL_OFF) {
    t.sub = A.tsub[i];
  } else if (KIND == HOR_P_OFF || KIND == HOR_SLDL_OFF) {
    t.sub = A.tsub[i];             // f0.n0
    t.g0x = A.tsub[ntp + i];       // P: n0 x f0; SLDL: g0
    t.g0y = A.tsub[2 * ntp + i];
    t.g0z = A.tsub[3 * ntp + i];
  } else {
    t.g0x = A.tsub[i];
    t.g0y = A.tsub[ntp + i];
    t.g0z = A.tsub[2 * ntp + i];
  }
  if (KIND == HOR_DL_OFF || KIND == HOR_SLDL_OFF) {
    t.x0x = A.tx0[i]; t.x0y = A.tx0[ntp + i]; t.x0z = A.tx0[2 * ntp + i];
    t.n0x = A.tn0[i]; t.n0y = A.tn0[ntp + i]; t.n0z = A.tn0[2 * ntp + i];
    t.b = A.tb[i];
  }
  if (NOSUB) t.g0x = t.g0y = t.g0z = 0.0;
}

// Sums over the sorted sources [j0, j1), which start at 64-subtile sub0.
// Source records are read with wave-uniform addresses (scalar loads).  tab is
// the kind's table in LDS (stage_table).
//
// Mixed precision (MIXED): a subtile beyond 8 delta of the wave runs the plain
// kernel in FP32, on coordinates relative to the subtile's box centre (REL32,
// the direct sums, whose FP32 records hold y - centre) or in FP64 (the leaf
// sums).  In a mixed subtile a per-source wave vote sends the source to the
// FP64 regularized code if some lane is within 4 delta, else to the FP32
// regularized code if some lane is within 8 delta (d from the FP64
// difference, so no cancellation), else to the FP64 plain kernel (which
// reuses the vote's FP64 difference).  FP32 contributions are summed in FP32
// over one subtile (at most 64 terms) and the subtile sums in FP64.  The
// on-surface stresslet keeps every regularized pair in FP64: its tabulated
// kernel costs less than the FP32 formula.
template <int KIND, bool MIXED, bool REL32>
__device__ __forceinline__ void range_sum(const DirectArgs& A, int j0, int j1, int sub0, const double* wb,
                                          bool all_far, const Tgt& t, const double* __restrict__ tab,
                                          double u[3]) {
  const double rp = 1.7724538509055160273;
  const double inv_del = 1.0 / A.DEL;
  const double inv_del2 = inv_del * inv_del;
  const double inv_del5 = inv_del2 * inv_del2 * inv_del;
  const double self1 = 32.0 / 5.0 / rp / A.DEL;
  const double self2 = 256.0 / 15.0 / rp / (A.DEL * A.DEL * A.DEL);
  const SrcRec* __restrict__ R = A.rec;
  const unsigned long long nearm = wave_mask(t.nearp);   // lanes whose target is near its surface
  const SrcRec32* __restrict__ R32 = A.rec32;
  constexpr bool SL = (KIND == HOR_SL_ON || KIND == HOR_SL_OFF);
  const double k5 = vreg(4.375);   // rsqrt5_64's e^2 coefficient, in vector registers for the loops
  Tgt32 v;
  if (MIXED) make_tgt32(t, v);
  const float inv_del32 = (float)inv_del;
  for (int base = j0, sb = sub0; base < j1; base += kWave, sb++) {
    const int end = min(j1, base + kWave);
    const double* box = A.sub_box + 6 * sb;
    const bool fast = all_far || boxes_far(wb, box, A.far_d2);
    if (fast) {
      if (MIXED && REL32) {
        const float cx = (float)(t.px - 0.5 * (box[0] + box[3]));
        const float cy = (float)(t.py - 0.5 * (box[1] + box[4]));
        const float cz = (float)(t.pz - 0.5 * (box[2] + box[5]));
        float w[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 2
        for (int j = base; j < end; j++) {
          const SrcRec32& s = R32[j];
          if (SL) sl_plain32(s, cx - s.x, cy - s.y, cz - s.z, v.sub, w);
          else dl_plain32(s, cx - s.x, cy - s.y, cz - s.z, v.g0x, v.g0y, v.g0z, w);
        }
        u[0] += (double)w[0];
        u[1] += (double)w[1];
        u[2] += (double)w[2];
      } else {
#pragma unroll 2
        for (int j = base; j < end; j++) {
          if (SL) sl_plain(R[j], t.px, t.py, t.pz, t.sub, u);
          else dl_plain(R[j], t.px, t.py, t.pz, t.g0x, t.g0y, t.g0z, k5, u);
        }
      }
    } else {
      // Mixed subtile: a wave vote per source decides whether any lane needs
      // the regularized kernel (r <= 8 delta, with a relative margin); when
      // none does, the source takes the plain path for the whole wave.
      float w[3] = {0.0f, 0.0f, 0.0f};
      for (int j = base; j < end; j++) {
        const SrcRec s = R[j];
        const double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const unsigned long long need = wave_mask(!(r2 > A.near_r2)) & nearm;
        if (MIXED && KIND != HOR_DL_ON && need && !(wave_mask(!(r2 > A.near64_r2)) & nearm)) {
          const SrcRec32& s32 = R32[j];
          const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
          if (SL) sl_reg32(s32, fx, fy, fz, v, t.nearp, inv_del32, w);
          else if (KIND == HOR_DL_OFF) dl_off_reg32(s32, fx, fy, fz, v, t.nearp, inv_del32, w);
          else dl_on_reg32(s32, fx, fy, fz, v, inv_del32, w);
        } else if (need) {
          if (SL) sl_full(s, t.px, t.py, t.pz, t.sub, t.nearp, t.q, inv_del, inv_del2, self1, self2, tab, u);
          else if (KIND == HOR_DL_OFF) dl_off_full(s, t, inv_del, inv_del2, self2, tab, u);
          else dl_on_full(s, t, inv_del2 * (kDlOnIntervals / kDlOnSmax), inv_del5, tab, k5, u);
        } else {
          if (SL) sl_plain(s, t.px, t.py, t.pz, t.sub, u);
          else dl_plain(s, t.px, t.py, t.pz, t.g0x, t.g0y, t.g0z, k5, u);
        }
      }
      if (MIXED) {
        u[0] += (double)w[0];
        u[1] += (double)w[1];
        u[2] += (double)w[2];
      }
    }
  }
}

// Regularized pressure over the sorted sources [j0, j1) (HOR_P_OFF, FP64):
// the subtile test and per-source vote of range_sum on the pressure records,
// in each subtile's frame (the records' positions are relative to its box
// centre).
__device__ __forceinline__ void range_sum_p(const DirectArgs& A, int j0, int j1, int sub0, const double* wb,
                                            bool all_far, const Tgt& t, const double* __restrict__ tab,
                                            double u[3]) {
  const double inv_del = 1.0 / A.DEL;
  const double inv_del2 = inv_del * inv_del;
  const SrcRecP* __restrict__ R = A.recp;
  const unsigned long long nearm = wave_mask(t.nearp);   // lanes whose target is near its surface
  const double k3 = vreg(1.875);   // rsqrt3_64's e^2 coefficient, in vector registers for the loops
  for (int base = j0, sb = sub0; base < j1; base += kWave, sb++) {
    const int end = min(j1, base + kWave);
    const double* box = A.sub_box + 6 * sb;
    TgtP v;
    make_tgtp(t, box, v);
    if (all_far || boxes_far(wb, box, A.far_d2)) {
#pragma unroll 2
      for (int j = base; j < end; j++) p_plain(R[j], t, v, k3, u);
    } else {
      for (int j = base; j < end; j++) {
        const SrcRecP s = R[j];
        const double dx = v.px - s.x, dy = v.py - s.y, dz = v.pz - s.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (wave_mask(!(r2 > A.near_r2)) & nearm) p_full(s, t, v, r2, inv_del, inv_del2, tab, u);
        else p_plain_r2(s, t, v, r2, k3, u);
      }
    }
  }
}

// Stokeslet + stresslet over the sorted sources [j0, j1) (HOR_SLDL_OFF): the
// subtile test, the per-source vote and the pair functions of the separate
// SL_OFF and DL_OFF sums (range_sum), in the same order.  MIXED: far subtiles
// in FP32 on subtile-centred coordinates; in a mixed subtile FP64 if some lane
// is within 4 delta, FP32 regularized if within 8 delta, FP64 plain otherwise;
// FP32 terms summed per subtile in FP32 and the subtiles in FP64.
template <bool MIXED>
__device__ __forceinline__ void range_sum_sldl(const DirectArgs& A, int j0, int j1, int sub0, const double* wb,
                                               bool all_far, const Tgt& t, const double* __restrict__ tab,
                                               double us[3], double ud[3]) {
  const double rp = 1.7724538509055160273;
  const double inv_del = 1.0 / A.DEL;
  const double inv_del2 = inv_del * inv_del;
  const double self1 = 32.0 / 5.0 / rp / A.DEL;
  const double self2 = 256.0 / 15.0 / rp / (A.DEL * A.DEL * A.DEL);
  const SrcRecV* __restrict__ R = A.recv;
  const unsigned long long nearm = wave_mask(t.nearp);   // lanes whose target is near its surface
  const SrcRecV32* __restrict__ R32 = A.recv32;
  Tgt32 v;
  if (MIXED) make_tgt32(t, v);
  const float inv_del32 = (float)inv_del;
  for (int base = j0, sb = sub0; base < j1; base += kWave, sb++) {
    const int end = min(j1, base + kWave);
    const double* box = A.sub_box + 6 * sb;
    if (all_far || boxes_far(wb, box, A.far_d2)) {
      if (MIXED) {
        const float cx = (float)(t.px - 0.5 * (box[0] + box[3]));
        const float cy = (float)(t.py - 0.5 * (box[1] + box[4]));
        const float cz = (float)(t.pz - 0.5 * (box[2] + box[5]));
        float ws[3] = {0.0f, 0.0f, 0.0f}, wd[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 2
        for (int j = base; j < end; j++) {
          const SrcRecV32& s = R32[j];
          sldl_plain32(s, cx - s.x, cy - s.y, cz - s.z, v, ws, wd);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
          us[k] += (double)ws[k];
          ud[k] += (double)wd[k];
        }
      } else {
#pragma unroll 2
        for (int j = base; j < end; j++) sldl_plain(R[j], t, us, ud);
      }
    } else {
      float ws[3] = {0.0f, 0.0f, 0.0f}, wd[3] = {0.0f, 0.0f, 0.0f};
      for (int j = base; j < end; j++) {
        const SrcRecV s = R[j];
        const double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const unsigned long long need = wave_mask(!(r2 > A.near_r2)) & nearm;
        if (MIXED && need && !(wave_mask(!(r2 > A.near64_r2)) & nearm))
          sldl_reg32(R32[j], (float)dx, (float)dy, (float)dz, v, t.nearp, inv_del32, ws, wd);
        else if (need) sldl_full(s, t, inv_del, inv_del2, self1, self2, tab, us, ud);
        else sldl_plain(s, t, us, ud);
      }
      if (MIXED) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          us[k] += (double)ws[k];
          ud[k] += (double)wd[k];
        }
      }
    }
  }
}

}  // namespace hor
