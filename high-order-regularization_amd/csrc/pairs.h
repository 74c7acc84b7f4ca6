// Pair functions shared by the direct-sum and treecode leaf kernels.
//
// Pair arithmetic follows the reference's per-pair expressions
// (highreg_stokes_2surf/utilities.cpp:179-323, stokeslets.cpp:106-130,
// stresslets.cpp:117-168, 244-269).
#pragma once

#include "device.h"
#include "direct.h"

namespace hor {

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
template <bool F32>
__device__ __forceinline__ void sl_plain(const SrcRec& s, const SrcRec32& s32, double px, double py,
                                         double pz, float pxf, float pyf, float pzf, double sub,
                                         double u[3]) {
  if (F32) {
    float dx = pxf - s32.x, dy = pyf - s32.y, dz = pzf - s32.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    float ri = rsqrt32(r2);
    float subf = (float)sub;
    float fx = s32.d0 - subf * s32.nx, fy = s32.d1 - subf * s32.ny, fz = s32.d2 - subf * s32.nz;
    float t = (fx * dx + fy * dy + fz * dz) * (ri * ri);
    u[0] += (double)((fx + dx * t) * ri);
    u[1] += (double)((fy + dy * t) * ri);
    u[2] += (double)((fz + dz * t) * ri);
  } else {
    double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    double ri = rsqrt64(r2);
    double fx = s.d0 - sub * s.nx, fy = s.d1 - sub * s.ny, fz = s.d2 - sub * s.nz;
    double t = (fx * dx + fy * dy + fz * dz) * (ri * ri);
    u[0] += (fx + dx * t) * ri;
    u[1] += (fy + dy * t) * ri;
    u[2] += (fz + dz * t) * ri;
  }
}

// ---- SL pair, full branch structure (mixed path), FP64
// (utilities.cpp:179-207, stokeslets.cpp:106-130).  r = sqrt(r2) and 1/r come
// from one reciprocal square root; the self test r < 1e-14 is r2 < 1e-28.
__device__ __forceinline__ void sl_full(const SrcRec& s, double px, double py, double pz,
                                        double sub, bool nearp, const double* q, double inv_del,
                                        double self1, double self2, const double* ec, double u[3]) {
  double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double fx = s.d0 - sub * s.nx, fy = s.d1 - sub * s.ny, fz = s.d2 - sub * s.nz;
  double fdx = fx * dx + fy * dy + fz * dz;
  double s1, s2;
  if (nearp && r2 < 1e-28) {
    s1 = self1;
    s2 = self2;
  } else {
    double ri = rsqrt64(r2);
    double rd = (r2 * ri) * inv_del;
    double ri3 = ri * ri * ri;
    if (!nearp || rd > 8.0) {
      s1 = ri;
      s2 = ri3;
    } else {
      double rd2 = rd * rd;
      double E = exp(-rd2);
      double ef = erf_from_exp(rd, E, ec);
      double rex = rd * E;
      double p1 = q[0] + rd2 * (q[1] + rd2 * q[2]);
      double p2 = q[3] + rd2 * (q[4] + rd2 * (q[5] + rd2 * q[6]));
      s1 = (ef + p1 * rex) * ri;
      s2 = (ef + p2 * rex) * ri3;
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
template <bool F32>
__device__ __forceinline__ void dl_plain(const SrcRec& s, const SrcRec32& s32, double px, double py,
                                         double pz, float pxf, float pyf, float pzf, double g0x,
                                         double g0y, double g0z, double u[3]) {
  if (F32) {
    float dx = pxf - s32.x, dy = pyf - s32.y, dz = pzf - s32.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    float ri = rsqrt32(r2);
    float ri2 = ri * ri;
    float gdx = (s32.d0 - (float)g0x) * dx + (s32.d1 - (float)g0y) * dy + (s32.d2 - (float)g0z) * dz;
    float ndx = s32.nx * dx + s32.ny * dy + s32.nz * dz;
    float k = gdx * ndx * (ri2 * ri2 * ri);
    u[0] += (double)(dx * k);
    u[1] += (double)(dy * k);
    u[2] += (double)(dz * k);
  } else {
    double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
    double r2 = dx * dx + dy * dy + dz * dz;
    double ri = rsqrt64(r2);
    double ri2 = ri * ri;
    double gdx = (s.d0 - g0x) * dx + (s.d1 - g0y) * dy + (s.d2 - g0z) * dz;
    double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
    double k = gdx * ndx * (ri2 * ri2 * ri);
    u[0] += dx * k;
    u[1] += dy * k;
    u[2] += dz * k;
  }
}

// ---- DL off-surface pair, full branch structure (stresslets.cpp:117-168).
// Every term is linear in the source normal, so the record's a n carries the
// weight.
__device__ __forceinline__ void dl_off_full(const SrcRec& s, const Tgt& t, double inv_del,
                                            double self2, const double* ec, double u[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  double gx = s.d0 - t.g0x, gy = s.d1 - t.g0y, gz = s.d2 - t.g0z;
  double H2, H3;
  if (t.nearp && r2 < 1e-28) {
    H2 = self2;
    H3 = 0.0;
  } else {
    double ri = rsqrt64(r2);
    double rd = (r2 * ri) * inv_del;
    double ri2 = ri * ri;
    if (!t.nearp || rd > 8.0) {
      double gdx = gx * dx + gy * dy + gz * dz;
      double ndx = s.nx * dx + s.ny * dy + s.nz * dz;
      double k = gdx * ndx * (ri2 * ri2 * ri);
      u[0] += dx * k;
      u[1] += dy * k;
      u[2] += dz * k;
      return;
    }
    double ri3 = ri2 * ri;
    double rd2 = rd * rd;
    double E = exp(-rd2);
    double ef = erf_from_exp(rd, E, ec);
    double rex = rd * E;
    const double* q = t.q;
    double p2 = q[3] + rd2 * (q[4] + rd2 * (q[5] + rd2 * q[6]));
    double p3 = q[7] + rd2 * (q[8] + rd2 * (q[9] + rd2 * (q[10] + rd2 * q[11])));
    H2 = (ef + p2 * rex) * ri3;
    H3 = (ef + p3 * rex) * (ri3 * ri2);
  }
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
  u[0] += dx * Bd - t.n0x * Bn + xh0 * Bx;
  u[1] += dy * Bd - t.n0y * Bn + xh1 * Bx;
  u[2] += dz * Bd - t.n0z * Bn + xh2 * Bx;
}

// ---- DL on-surface pair, full branch structure (stresslets.cpp:244-269).
__device__ __forceinline__ void dl_on_full(const SrcRec& s, const Tgt& t, double inv_del,
                                           const double* ec, double u[3]) {
  double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
  double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 < 1e-28) return;  // d = 0: the self term contributes exactly zero
  double ri = rsqrt64(r2);
  double rd = (r2 * ri) * inv_del;
  double ri2 = ri * ri;
  double H3;
  if (rd > 8.0) {
    H3 = ri2 * ri2 * ri;
  } else {
    double rd2 = rd * rd;
    double E = exp(-rd2);
    double ef = erf_from_exp(rd, E, ec);
    double rex = rd * E;
    const double* q = t.q;
    double p3 = q[7] + rd2 * (q[8] + rd2 * (q[9] + rd2 * q[10]));
    H3 = (ef + p3 * rex) * (ri2 * ri2 * ri);
  }
  double gx = s.d0 - t.g0x, gy = s.d1 - t.g0y, gz = s.d2 - t.g0z;
  double dxn = dx * s.nx + dy * s.ny + dz * s.nz;
  double dxg = dx * gx + dy * gy + dz * gz;
  H3 *= dxn * dxg;
  u[0] += dx * H3;
  u[1] += dy * H3;
  u[2] += dz * H3;
}

template <int KIND>
__device__ __forceinline__ void load_target(const DirectArgs& A, int i, Tgt& t) {
  int ntp = A.ntp;
  t.px = A.tx[i];
  t.py = A.tx[ntp + i];
  t.pz = A.tx[2 * ntp + i];
  t.nearp = A.tnear[i] != 0;
#pragma unroll
  for (int k = 0; k < kNumCoef; k++) t.q[k] = A.tcoef[k * ntp + i];
  t.sub = 0;
  t.g0x = t.g0y = t.g0z = 0;
  t.x0x = t.x0y = t.x0z = t.n0x = t.n0y = t.n0z = t.b = 0;
  if (KIND == HOR_SL_ON || KIND == HOR_SL_OFF) {
    t.sub = A.tsub[i];
  } else {
    t.g0x = A.tsub[i];
    t.g0y = A.tsub[ntp + i];
    t.g0z = A.tsub[2 * ntp + i];
  }
  if (KIND == HOR_DL_OFF) {
    t.x0x = A.tx0[i]; t.x0y = A.tx0[ntp + i]; t.x0z = A.tx0[2 * ntp + i];
    t.n0x = A.tn0[i]; t.n0y = A.tn0[ntp + i]; t.n0z = A.tn0[2 * ntp + i];
    t.b = A.tb[i];
  }
}

// Sums over the sorted sources [j0, j1), which start at 64-subtile sub0.
// Source records are read with wave-uniform addresses (scalar loads).
template <int KIND, bool MIXED>
__device__ __forceinline__ void range_sum(const DirectArgs& A, int j0, int j1, int sub0, const double* wb,
                                          bool all_far, const Tgt& t, float pxf, float pyf, float pzf,
                                          double u[3]) {
  const double rp = 1.7724538509055160273;
  const double inv_del = 1.0 / A.DEL;
  const double self1 = 32.0 / 5.0 / rp / A.DEL;
  const double self2 = 256.0 / 15.0 / rp / (A.DEL * A.DEL * A.DEL);
  const SrcRec* __restrict__ R = A.rec;
  const SrcRec32* __restrict__ R32 = A.rec32;
  for (int base = j0, sb = sub0; base < j1; base += kWave, sb++) {
    const int end = min(j1, base + kWave);
    const bool fast = all_far || boxes_far(wb, A.sub_box + 6 * sb, A.far_d2);
    if (fast) {
#pragma unroll 2
      for (int j = base; j < end; j++) {
        if (KIND == HOR_SL_ON || KIND == HOR_SL_OFF)
          sl_plain<MIXED>(R[j], MIXED ? R32[j] : SrcRec32{}, t.px, t.py, t.pz, pxf, pyf, pzf, t.sub, u);
        else
          dl_plain<MIXED>(R[j], MIXED ? R32[j] : SrcRec32{}, t.px, t.py, t.pz, pxf, pyf, pzf, t.g0x, t.g0y,
                          t.g0z, u);
      }
    } else {
      // Mixed subtile: a wave vote per source decides whether any lane needs
      // the regularized kernel (r <= 8 delta, with a relative margin); when
      // none does, the source takes the plain path for the whole wave.
      for (int j = base; j < end; j++) {
        const SrcRec s = R[j];
        const double dx = t.px - s.x, dy = t.py - s.y, dz = t.pz - s.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const bool need = t.nearp && !(r2 > A.near_r2);
        if (__any(need ? 1 : 0)) {
          if (KIND == HOR_SL_ON || KIND == HOR_SL_OFF)
            sl_full(s, t.px, t.py, t.pz, t.sub, t.nearp, t.q, inv_del, self1, self2, A.erfcx, u);
          else if (KIND == HOR_DL_OFF)
            dl_off_full(s, t, inv_del, self2, A.erfcx, u);
          else
            dl_on_full(s, t, inv_del, A.erfcx, u);
        } else {
          if (KIND == HOR_SL_ON || KIND == HOR_SL_OFF)
            sl_plain<false>(s, SrcRec32{}, t.px, t.py, t.pz, pxf, pyf, pzf, t.sub, u);
          else
            dl_plain<false>(s, SrcRec32{}, t.px, t.py, t.pz, pxf, pyf, pzf, t.g0x, t.g0y, t.g0z, u);
        }
      }
    }
  }
}

}  // namespace hor
