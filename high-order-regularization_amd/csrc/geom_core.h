// Level-set geometry of an ellipsoid body, shared by the host geometry
// (geometry.cpp, built by g++) and its GPU counterpart (geom_gpu.hip, built
// by hipcc), so that both execute the same IEEE operations in the same order.
// The host build has no fused multiply-add instructions; translation units
// that compile this header for the device turn contraction off
// (-ffp-contract=off), so the two agree bit for bit wherever only +, -, *,
// /, sqrt and comparisons are involved.  bump() (exp, acos) is the one place
// where the two math libraries may differ, by an ulp or two.
#pragma once

#include <cmath>

#include "hor.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HOR_GEOM __host__ __device__ inline
#else
#define HOR_GEOM inline
#endif

namespace hor {
namespace geom {

// Level set of an ellipsoid body, x_body = R (x - ctr)
// (highreg_stokes_2surf/utilities.cpp:44-158).

HOR_GEOM void to_body(const hor_body& B, const double x[3], bool shift, double u[3]) {
  double xx[3];
  for (int k = 0; k < 3; k++) xx[k] = shift ? x[k] - B.ctr[k] : x[k];
  if (B.rotated) {
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 3; j++) s += B.R[3 * i + j] * xx[j];
      u[i] = s;
    }
  } else {
    u[0] = xx[0]; u[1] = xx[1]; u[2] = xx[2];
  }
}

// phi (utilities.cpp:58-86); shift=false evaluates in the body-centred frame.
HOR_GEOM double phi(const hor_body& B, const double x[3], bool shift) {
  double u[3];
  to_body(B, x, shift, u);
  double a2[3] = {B.abc[0] * B.abc[0], B.abc[1] * B.abc[1], B.abc[2] * B.abc[2]};
  return u[0] * u[0] / a2[0] + u[1] * u[1] / a2[1] + u[2] * u[2] / a2[2] - 1.0;
}

// Dphi (utilities.cpp:91-121).
HOR_GEOM double dphi(const hor_body& B, int i, const double x[3], bool shift) {
  double a2[3] = {B.abc[0] * B.abc[0], B.abc[1] * B.abc[1], B.abc[2] * B.abc[2]};
  if (B.rotated) {
    double u[3];
    to_body(B, x, shift, u);
    double g[3] = {2.0 * u[0] / a2[0], 2.0 * u[1] / a2[1], 2.0 * u[2] / a2[2]};
    return g[0] * B.R[0 * 3 + i] + g[1] * B.R[1 * 3 + i] + g[2] * B.R[2 * 3 + i];
  }
  double xi = shift ? x[i] - B.ctr[i] : x[i];
  return 2.0 * xi / a2[i];
}

// D2phi (utilities.cpp:126-158): constant Hessian.
HOR_GEOM void d2phi(const hor_body& B, double H[9]) {
  double a2[3] = {B.abc[0] * B.abc[0], B.abc[1] * B.abc[1], B.abc[2] * B.abc[2]};
  if (B.rotated) {
    double A[3] = {2.0 / a2[0], 2.0 / a2[1], 2.0 / a2[2]};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        H[3 * i + j] = A[0] * B.R[i] * B.R[j] + A[1] * B.R[3 + i] * B.R[3 + j] +
                       A[2] * B.R[6 + i] * B.R[6 + j];
    H[3] = H[1]; H[6] = H[2]; H[7] = H[5];
  } else {
    for (int k = 0; k < 9; k++) H[k] = 0.0;
    H[0] = 2.0 / a2[0]; H[4] = 2.0 / a2[1]; H[8] = 2.0 / a2[2];
  }
}

HOR_GEOM int sign_phi(const hor_body& B, const double x[3]) {
  return phi(B, x, false) < 0.0 ? -1 : 1;  // surface.cpp:76-81
}

constexpr double kTol = 1e-14;  // highreg_stokes_2surf/utilities.h:14

// bisection (surface.cpp:139-174), on the segment PT + t*H e_i, t in [a,b].
HOR_GEOM double bisection(const hor_body& B, const double PT[3], double H, int i, double a, double b) {
  double xl[3] = {PT[0], PT[1], PT[2]}, xr[3] = {PT[0], PT[1], PT[2]}, xm[3] = {PT[0], PT[1], PT[2]};
  xl[i] += a * H;
  xr[i] += b * H;
  double p = 0.0;
  if (phi(B, xl, false) * phi(B, xr, false) > 0.0) return p;
  if (std::fabs(phi(B, xl, false)) <= kTol) return a;
  if (std::fabs(phi(B, xr, false)) <= kTol) return b;
  p = (a + b) / 2;
  xm[i] = PT[i] + p * H;
  double pm = phi(B, xm, false);
  int guard = 0;
  while (std::fabs(pm) > kTol && guard++ < 200) {
    if (phi(B, xl, false) * pm < 0.0) b = p;
    else { a = p; xl[0] = xm[0]; xl[1] = xm[1]; xl[2] = xm[2]; }
    p = (a + b) / 2;
    xm[i] = PT[i] + p * H;
    pm = phi(B, xm, false);
  }
  return p;
}

// Newton (surface.cpp:95-135) with the bisection fallback.
HOR_GEOM double newton(const hor_body& B, const double PT[3], double H, int i, double a, double b) {
  const int n_max = 100;
  double xl[3] = {PT[0], PT[1], PT[2]}, xr[3] = {PT[0], PT[1], PT[2]}, xn[3] = {PT[0], PT[1], PT[2]};
  xl[i] += a * H;
  xr[i] += b * H;
  double p = 0.0;
  if (phi(B, xl, false) * phi(B, xr, false) > 0.0) return p;
  if (std::fabs(phi(B, xl, false)) <= kTol) return a;
  if (std::fabs(phi(B, xr, false)) <= kTol) return b;
  p = (a + b) / 2;
  xn[i] = PT[i] + p * H;
  double pn = phi(B, xn, false);
  double dn = dphi(B, i, xn, false) * H;
  int n = 0;
  while ((std::fabs(pn) > kTol) && (std::fabs(dn) > kTol) && (n < n_max)) {
    p -= pn / dn;
    xn[i] = PT[i] + p * H;
    n = n + 1;
    pn = phi(B, xn, false);
    dn = dphi(B, i, xn, false) * H;
  }
  if ((n >= n_max) || (std::fabs(dn) <= kTol)) p = bisection(B, PT, H, i, a, b);
  return p;
}

// bump and Part_Unity (surface.cpp:181-207).
HOR_GEOM double bump(double r) {
  if (std::fabs(r) < 1.0) {
    double r2 = r * r;
    return std::exp(2.0 * r2 / (r2 - 1.0));
  }
  return 0.0;
}

HOR_GEOM double part_unity(int i, const double n[3], double theta) {
  double v0 = bump(std::acos(std::fabs(n[0])) / theta);
  double v1 = bump(std::acos(std::fabs(n[1])) / theta);
  double v2 = bump(std::acos(std::fabs(n[2])) / theta);
  double v[3] = {v0, v1, v2};
  double s = v0 + v1 + v2;
  return v[i] / s;
}

// ---------------------------------------------------------------------------
// 4x4 LU with partial pivoting (utilities.cpp:361-431).

HOR_GEOM int lu4(double A[4][4], int p[4]) {
  const int n = 4;
  for (int j = 0; j < n; j++) p[j] = j;
  for (int j = 0; j < n; j++) {
    int k = j;
    double m = A[p[j]][j];
    for (int i = j + 1; i < n; i++)
      if (std::fabs(A[p[i]][j]) > std::fabs(m)) { m = A[p[i]][j]; k = i; }
    if (k != j) { int t = p[j]; p[j] = p[k]; p[k] = t; }
    double ajj = A[p[j]][j];
    if (std::fabs(ajj) < 1.0e-15) return -1;
    for (int i = j + 1; i < n; i++) {
      double lij = A[p[i]][j] / ajj;
      A[p[i]][j] = lij;
      for (int kk = j + 1; kk < n; kk++) A[p[i]][kk] -= lij * A[p[j]][kk];
    }
  }
  return 0;
}

HOR_GEOM void lu4_solve(double A[4][4], const int p[4], const double b[4], double x[4]) {
  const int n = 4;
  x[0] = b[p[0]];
  for (int i = 1; i < n; i++) {
    x[i] = b[p[i]];
    double s = 0.0;
    for (int j = 0; j < i; j++) s += A[p[i]][j] * x[j];
    x[i] -= s;
  }
  x[n - 1] = x[n - 1] / A[p[n - 1]][n - 1];
  for (int i = n - 2; i >= 0; i--) {
    double s = 0.0;
    for (int j = n - 1; j > i; j--) s += A[p[i]][j] * x[j];
    x[i] = (x[i] - s) / A[p[i]][i];
  }
}

// find_Nearest_Start (surface.cpp:277-365): Newton on (lambda, x) for the
// closest point of the level set to pt, started from xs.
HOR_GEOM int nearest_point(const hor_body& B, const double pt[3], const double xs[3], double x0[3],
                         double n0[3]) {
  const int n_max = 100;
  double x[3] = {xs[0], xs[1], xs[2]};
  double lam = 0.0;
  int n = 0;
  double F[4];
  F[0] = phi(B, x, true);
  F[1] = x[0] - pt[0];
  F[2] = x[1] - pt[1];
  F[3] = x[2] - pt[2];
  double Fn2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
  double H[9];
  d2phi(B, H);
  while ((Fn2 > kTol) && (n <= n_max)) {
    double g[3] = {dphi(B, 0, x, true), dphi(B, 1, x, true), dphi(B, 2, x, true)};
    F[0] = phi(B, x, true);
    F[1] = x[0] - pt[0] + lam * g[0];
    F[2] = x[1] - pt[1] + lam * g[1];
    F[3] = x[2] - pt[2] + lam * g[2];
    Fn2 = F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + F[3] * F[3];
    double J[4][4] = {{0.0, g[0], g[1], g[2]},
                      {g[0], 1.0 + lam * H[0], lam * H[1], lam * H[2]},
                      {g[1], lam * H[3], 1.0 + lam * H[4], lam * H[5]},
                      {g[2], lam * H[6], lam * H[7], 1.0 + lam * H[8]}};
    int p[4];
    if (lu4(J, p) == -1) return -1;
    double s[4];
    lu4_solve(J, p, F, s);
    lam = lam - s[0];
    for (int i = 0; i < 3; i++) x[i] = x[i] - s[i + 1];
    n = n + 1;
  }
  if (n > n_max) return -1;
  double g[3] = {dphi(B, 0, x, true), dphi(B, 1, x, true), dphi(B, 2, x, true)};
  double nn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  for (int k = 0; k < 3; k++) {
    x0[k] = x[k];
    n0[k] = g[k] / nn;
  }
  return 0;
}

// select_Monge_Patch (surface.cpp:395-431).
HOR_GEOM void monge_patch(const hor_body& B, const double x[3], int& i1, int& i2, int& i3) {
  double nx = std::fabs(dphi(B, 0, x, true));
  double ny = std::fabs(dphi(B, 1, x, true));
  double nz = std::fabs(dphi(B, 2, x, true));
  if (nx >= ny) {
    if (nx >= nz) { i1 = 1; i2 = 2; i3 = 0; }
    else { i1 = 0; i2 = 1; i3 = 2; }
  } else {
    if (ny >= nz) { i1 = 0; i2 = 2; i3 = 1; }
    else { i1 = 0; i2 = 1; i3 = 2; }
  }
}

// Mean curvature (surface.cpp:214-272): the reference's spheroid formula when
// the body is a spheroid about its first axis (b == c), else the general
// level-set formula that follows it in the same function.
HOR_GEOM double mean_curvature(const hor_body& B, const double x[3]) {
  if (B.abc[1] == B.abc[2]) {
    double a = B.abc[1], c = B.abc[0];
    double a2 = a * a;
    double u[3];
    to_body(B, x, true, u);
    double cos2 = (u[1] * u[1] + u[2] * u[2]) / a2;
    double t = a2 + (c * c - a2) * cos2;
    return c * (a2 + t) / (2.0 * a * std::sqrt(t * t * t));
  }
  double p1 = dphi(B, 0, x, true), p2 = dphi(B, 1, x, true), p3 = dphi(B, 2, x, true);
  double H[9];
  d2phi(B, H);
  double p11 = H[0], p12 = H[1], p13 = H[2], p22 = H[4], p23 = H[5], p33 = H[8];
  double s1 = p1 * p1, s2 = p2 * p2, s3 = p3 * p3;
  double r = s1 * p22 - 2.0 * p1 * p2 * p12 + s2 * p11;
  r += s1 * p33 - 2.0 * p1 * p3 * p13 + s3 * p11;
  r += s2 * p33 - 2.0 * p2 * p3 * p23 + s3 * p22;
  double d = s1 + s2 + s3;
  return r / (2.0 * d * std::sqrt(d));
}

// surfc_force (surface.cpp:530-559).
HOR_GEOM void surface_force(const hor_body& B, const double x0[3], const double n0[3], double f[3]) {
  double H = mean_curvature(B, x0);
  double zz = x0[2] - B.ctr[2];
  double gamma = 1.0 + zz * zz;
  double dgx = 0.0, dgy = 0.0, dgz = 2.0 * zz;
  double dgn = dgx * n0[0] + dgy * n0[1] + dgz * n0[2];
  double sg0 = dgx - dgn * n0[0], sg1 = dgy - dgn * n0[1], sg2 = dgz - dgn * n0[2];
  gamma *= 2.0 * H;
  f[0] = gamma * n0[0] - sg0;
  f[1] = gamma * n0[1] - sg1;
  f[2] = gamma * n0[2] - sg2;
}

// sixPoint_coeff (main.cpp:750-757): 6-point Lagrange weights at offset p for
// the nodes -2..3.
HOR_GEOM void six_coeff(double p, double a[6]) {
  a[0] = -p * (p * p - 1.0) * (p - 2.0) * (p - 3.0) / 120.0;
  a[1] = p * (p - 1.0) * (p * p - 4.0) * (p - 3.0) / 24.0;
  a[2] = -(p * p - 1.0) * (p * p - 4.0) * (p - 3.0) / 12.0;
  a[3] = p * (p + 1.0) * (p * p - 4.0) * (p - 3.0) / 12.0;
  a[4] = -p * (p * p - 1.0) * (p + 2.0) * (p - 3.0) / 24.0;
  a[5] = p * (p * p - 1.0) * (p * p - 4.0) / 120.0;
}

}  // namespace geom
}  // namespace hor
