/* CPU oracle for the regularized Stokes boundary-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see hor_oracle.h): a plain-C restatement of the
 * reference algorithms, used to check the HIP path.  Every function cites the
 * reference file:line it follows (paths relative to the reference root).
 * Per-pair arithmetic follows the reference's expression order so that the
 * oracle agrees with the compiled reference to a few ulps.
 */
#include "hor_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

static double rootpi(void) { return sqrt(ORC_PI); }

/* ======================================================================
 * Scalar kernels
 * ====================================================================== */

/* evaluate_In_024 (highreg_stokes_2surf/utilities.cpp:327-337):
 * I_{2k}(lambda) = int_{|lambda|}^inf (s^2 - lambda^2)^k erfc(s) ds, k = 0,1,2,
 * in closed form. */
void orc_In_024(double lam, double I[3]) {
  const double rp = rootpi();
  double lam2 = lam * lam;
  double em2 = exp(-lam2);
  double t = fabs(lam) * erfc(fabs(lam));
  I[0] = em2 / rp - t;
  I[1] = 2.0 / 3.0 * ((0.5 - lam2) * em2 / rp + lam2 * t);
  I[2] = 8.0 / 15.0 * ((0.75 - 0.5 * lam2 + lam2 * lam2) * em2 / rp - lam2 * lam2 * t);
}

/* get_reg_coeff_7ord (utilities.cpp:163-175): lambda = b/delta -> c1,c2,c3. */
void orc_reg_coeff(double lam, double c[3]) {
  double lam2 = lam * lam;
  double I[3];
  orc_In_024(lam, I);
  double I0 = I[0], I2 = I[1], I4 = I[2];
  double e2 = exp(lam2);
  double c3 = e2 * (2.0 * I0 - 4.0 * I2 + I4) / 8.0;
  double c2 = e2 * (I0 - I2) + (4.0 * lam2 + 7.0) * c3;
  double c1 = 2.0 * e2 * I0 + 2.0 * (lam2 + 1.0) * c2 - (4.0 * lam2 * lam2 + 6.0 * lam2 + 6.0) * c3;
  c[0] = c1;
  c[1] = c2;
  c[2] = c3;
}

/* Stokes_SL_7ord (utilities.cpp:179-207). */
void orc_sl_kernel(double r, double d, const double c[3], double* s1, double* s2) {
  const double rp = rootpi();
  if (r < 1e-14) {
    *s1 = 32.0 / 5.0 / rp / d;
    *s2 = 256.0 / 15.0 / rp / (d * d * d);
  } else if (r / d > 8.0) {
    double r2 = r * r;
    *s1 = 1.0 / r;
    *s2 = 1.0 / (r2 * r);
  } else {
    double rd = r / d, rd2 = rd * rd, rd4 = rd2 * rd2, rd6 = rd4 * rd2;
    double ef = erf(rd);
    double rex = rd * exp(-rd2);
    double p1 = c[0] - 2.0 * (c[1] + c[2]) * rd2 + 4.0 * c[2] * rd4;
    double p2 = -2.0 / rp + 2.0 * (c[0] + 2.0 * (c[1] + c[2])) * rd2
                - 4.0 * (c[1] + 5.0 * c[2]) * rd4 + 8.0 * c[2] * rd6;
    double r2 = r * r;
    *s1 = (ef + p1 * rex) / r;
    *s2 = (ef + p2 * rex) / (r2 * r);
  }
}

/* Stokes_SL_7ord_onSurf (utilities.cpp:211-254), 7th-order branch. */
void orc_sl_kernel_on(double r, double d, double* s1, double* s2) {
  const double rp = rootpi();
  if (r < 1e-14) {
    *s1 = 32.0 / 5.0 / rp / d;
    *s2 = 256.0 / 15.0 / rp / (d * d * d);
  } else if (r / d > 8.0) {
    double r2 = r * r;
    *s1 = 1.0 / r;
    *s2 = 1.0 / (r2 * r);
  } else {
    double rd = r / d, rd2 = rd * rd, rd4 = rd2 * rd2, rd6 = rd4 * rd2;
    double ef = erf(rd);
    double rex = 2.0 / rp * rd * exp(-rd2);
    double p1 = 11.0 / 5.0 - 26.0 / 15.0 * rd2 + 4.0 / 15.0 * rd4;
    double p2 = -1.0 + (118.0 * rd2 - 68.0 * rd4 + 8.0 * rd6) / 15.0;
    double r2 = r * r;
    *s1 = (ef + p1 * rex) / r;
    *s2 = (ef + p2 * rex) / (r2 * r);
  }
}

/* Stokes_DL_7ord (utilities.cpp:258-288). */
void orc_dl_kernel(double r, double d, const double c[3], double* s2, double* s3) {
  const double rp = rootpi();
  if (r < 1e-14) {
    *s2 = 256.0 / 15.0 / rp / (d * d * d);
    *s3 = 0.0;
  } else if (r / d > 8.0) {
    double r2 = r * r;
    *s2 = 1.0 / (r2 * r);
    *s3 = 1.0 / (r2 * r2 * r);
  } else {
    double rd = r / d, rd2 = rd * rd, rd4 = rd2 * rd2, rd6 = rd4 * rd2;
    double ef = erf(rd);
    double rex = rd * exp(-rd2);
    double p2 = -2.0 / rp + 2.0 * (c[0] + 2.0 * (c[1] + c[2])) * rd2
                - 4.0 * (c[1] + 5.0 * c[2]) * rd4 + 8.0 * c[2] * rd6;
    double p3 = -2.0 / rp * (1.0 + 2.0 / 3.0 * rd2)
                + 4.0 / 3.0 * ((c[0] + 4.0 * c[1] + 12.0 * c[2]) * rd4
                               - 2.0 * (c[1] + 9.0 * c[2]) * rd6 + 4.0 * c[2] * rd6 * rd2);
    double r2 = r * r;
    *s2 = (ef + p2 * rex) / (r2 * r);
    *s3 = (ef + p3 * rex) / (r2 * r2 * r);
  }
}

/* Stokes_DL_7ord_onSurf (utilities.cpp:292-323), 7th-order branch. */
void orc_dl_kernel_on(double r, double d, double* s3) {
  const double rp = rootpi();
  if (r < 1e-14) {
    *s3 = 128.0 / 45.0 / rp / (d * d * d * d * d);
  } else if (r / d > 8.0) {
    double r2 = r * r;
    *s3 = 1.0 / (r2 * r2 * r);
  } else {
    double rd = r / d, rd2 = rd * rd, rd4 = rd2 * rd2, rd6 = rd4 * rd2;
    double ef = erf(rd);
    double rex = 2.0 / rp * rd * exp(-rd2);
    double p3 = -1.0 - 2.0 / 3.0 * rd2 + 52.0 / 45.0 * rd4 - 8.0 / 45.0 * rd6;
    double r2 = r * r;
    *s3 = (ef + p3 * rex) / (r2 * r2 * r);
  }
}

/* ======================================================================
 * Per-target sums over a source range [j0, j1] (inclusive), unscaled.
 * These are the loop bodies shared by the direct sums and the treecode's
 * leaf sums (KITC.cpp:807-1062 repeat the same per-pair arithmetic).
 * ====================================================================== */

/* Plain Stokeslet (stokeslets.cpp:44-65; KITC.cpp:858-881). */
static void sum_sl_plain(int j0, int j1, const double* sx, const double* sa,
                         const double* sf, const double p[3], double u[3]) {
  for (int j = j0; j <= j1; j++) {
    double dx = p[0] - sx[3 * j], dy = p[1] - sx[3 * j + 1], dz = p[2] - sx[3 * j + 2];
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double fx = sf[3 * j], fy = sf[3 * j + 1], fz = sf[3 * j + 2];
    double fdx = fx * dx + fy * dy + fz * dz;
    double a = sa[j];
    double H1 = a / r;
    double H2 = fdx * a / (r2 * r);
    u[0] += fx * H1 + dx * H2;
    u[1] += fy * H1 + dy * H2;
    u[2] += fz * H1 + dz * H2;
  }
}

/* Regularized Stokeslet with subtraction f - (f0.n0) n (stokeslets.cpp:106-130,
 * 196-220; KITC.cpp:816-841, 889-914).  on != 0 uses the lambda=0 kernel. */
static void sum_sl_reg(int on, int j0, int j1, const double* sx, const double* sn,
                       const double* sa, const double* sf, double f0n0,
                       const double c[3], double DEL, const double p[3], double u[3]) {
  for (int j = j0; j <= j1; j++) {
    double dx = p[0] - sx[3 * j], dy = p[1] - sx[3 * j + 1], dz = p[2] - sx[3 * j + 2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double a = sa[j];
    double fx = sf[3 * j] - f0n0 * sn[3 * j];
    double fy = sf[3 * j + 1] - f0n0 * sn[3 * j + 1];
    double fz = sf[3 * j + 2] - f0n0 * sn[3 * j + 2];
    double fdx = fx * dx + fy * dy + fz * dz;
    double H1, H2;
    if (on)
      orc_sl_kernel_on(r, DEL, &H1, &H2);
    else
      orc_sl_kernel(r, DEL, c, &H1, &H2);
    H1 *= a;
    H2 *= a * fdx;
    u[0] += fx * H1 + dx * H2;
    u[1] += fy * H1 + dy * H2;
    u[2] += fz * H1 + dz * H2;
  }
}

/* Plain stresslet (stresslets.cpp:45-64; KITC.cpp:972-994). */
static void sum_dl_plain(int j0, int j1, const double* sx, const double* sn,
                         const double* sa, const double* sg, const double p[3],
                         double u[3]) {
  for (int j = j0; j <= j1; j++) {
    double dx = p[0] - sx[3 * j], dy = p[1] - sx[3 * j + 1], dz = p[2] - sx[3 * j + 2];
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double gdx = sg[3 * j] * dx + sg[3 * j + 1] * dy + sg[3 * j + 2] * dz;
    double ndx = sn[3 * j] * dx + sn[3 * j + 1] * dy + sn[3 * j + 2] * dz;
    double k = gdx * ndx / (r2 * r2 * r) * sa[j];
    u[0] += dx * k;
    u[1] += dy * k;
    u[2] += dz * k;
  }
}

/* Regularized stresslet off the surface with subtraction g - g0 and the
 * x_hat correction terms (stresslets.cpp:117-168; KITC.cpp:1002-1058). */
static void sum_dl_reg(int j0, int j1, const double* sx, const double* sn,
                       const double* sa, const double* sg, const double g0[3],
                       const double x0[3], const double n0[3], double b,
                       const double c[3], double DEL, const double p[3], double u[3]) {
  for (int j = j0; j <= j1; j++) {
    double dx = p[0] - sx[3 * j], dy = p[1] - sx[3 * j + 1], dz = p[2] - sx[3 * j + 2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double a = sa[j];
    double H2, H3;
    orc_dl_kernel(r, DEL, c, &H2, &H3);
    H2 *= a;
    H3 *= a;
    double xh0 = sx[3 * j] - x0[0], xh1 = sx[3 * j + 1] - x0[1], xh2 = sx[3 * j + 2] - x0[2];
    double sig2 = (xh0 * xh0 + xh1 * xh1 + xh2 * xh2) - 2.0 * b * (xh0 * n0[0] + xh1 * n0[1] + xh2 * n0[2]);
    double gx = sg[3 * j] - g0[0], gy = sg[3 * j + 1] - g0[1], gz = sg[3 * j + 2] - g0[2];
    double nx = sn[3 * j], ny = sn[3 * j + 1], nz = sn[3 * j + 2];
    double g_n0 = gx * n0[0] + gy * n0[1] + gz * n0[2];
    double g_xh = gx * xh0 + gy * xh1 + gz * xh2;
    double n0_n = n0[0] * nx + n0[1] * ny + n0[2] * nz;
    double xh_n = xh0 * nx + xh1 * ny + xh2 * nz;
    double m1a = g_n0 * n0_n;
    double m1b = g_xh * n0_n + g_n0 * xh_n;
    double m2a = g_xh * xh_n;
    double m2b = b * (g_xh * n0_n + g_n0 * xh_n);
    double t1x = dx * m1a - n0[0] * m1b;
    double t1y = dy * m1a - n0[1] * m1b;
    double t1z = dz * m1a - n0[2] * m1b;
    u[0] += t1x * H2;
    u[1] += t1y * H2;
    u[2] += t1z * H2;
    u[0] += (-sig2 * t1x + dx * m2a + xh0 * m2b) * H3;
    u[1] += (-sig2 * t1y + dy * m2a + xh1 * m2b) * H3;
    u[2] += (-sig2 * t1z + dz * m2a + xh2 * m2b) * H3;
  }
}

/* On-surface regularized stresslet (stresslets.cpp:244-269; KITC.cpp:930-954). */
static void sum_dl_on(int j0, int j1, const double* sx, const double* sn,
                      const double* sa, const double* sg, const double g0[3],
                      double DEL, const double p[3], double u[3]) {
  for (int j = j0; j <= j1; j++) {
    double dx = p[0] - sx[3 * j], dy = p[1] - sx[3 * j + 1], dz = p[2] - sx[3 * j + 2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double a = sa[j];
    double H3;
    orc_dl_kernel_on(r, DEL, &H3);
    double gx = sg[3 * j] - g0[0], gy = sg[3 * j + 1] - g0[1], gz = sg[3 * j + 2] - g0[2];
    double dxn = dx * sn[3 * j] + dy * sn[3 * j + 1] + dz * sn[3 * j + 2];
    double dxg = dx * gx + dy * gy + dz * gz;
    H3 *= dxn * dxg * a;
    u[0] += dx * H3;
    u[1] += dy * H3;
    u[2] += dz * H3;
  }
}

/* ======================================================================
 * Direct sums
 * ====================================================================== */

/* eval_Stokes_SL_offSurf + stokes_SL_7ord (stokeslets.cpp:12-135). */
void orc_sl_off(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sf, int nt, const double* tx,
                const double* tb, const double* tn0, const double* tf0, double* out) {
  const double s8 = 1.0 / (8.0 * ORC_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double p[3] = {tx[3 * i], tx[3 * i + 1], tx[3 * i + 2]};
    double b = tb[i];
    double u[3] = {0, 0, 0};
    if (fabs(b / DEL) > 8.0) {
      sum_sl_plain(0, ns - 1, sx, sa, sf, p, u);
    } else {
      double f0n0 = tf0[3 * i] * tn0[3 * i] + tf0[3 * i + 1] * tn0[3 * i + 1] + tf0[3 * i + 2] * tn0[3 * i + 2];
      double c[3];
      orc_reg_coeff(b / DEL, c);
      sum_sl_reg(0, 0, ns - 1, sx, sn, sa, sf, f0n0, c, DEL, p, u);
    }
    for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s8;
  }
}

/* eval_Stokes_SL_onSurf + stokes_SL_7ord_onSurf (stokeslets.cpp:141-225);
 * target i takes f0.n0 from source i (stokeslets.cpp:167). */
void orc_sl_on(double DEL, int ns, const double* sx, const double* sn,
               const double* sa, const double* sf, int nt, const double* tx,
               double* out) {
  const double s8 = 1.0 / (8.0 * ORC_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double p[3] = {tx[3 * i], tx[3 * i + 1], tx[3 * i + 2]};
    double f0n0 = sf[3 * i] * sn[3 * i] + sf[3 * i + 1] * sn[3 * i + 1] + sf[3 * i + 2] * sn[3 * i + 2];
    double u[3] = {0, 0, 0};
    sum_sl_reg(1, 0, ns - 1, sx, sn, sa, sf, f0n0, NULL, DEL, p, u);
    for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s8;
  }
}

/* eval_Stokes_DL_offSurf + stokes_DL_7ord (stresslets.cpp:13-179).  The jump
 * g0 is added only in the regularized branch, for b<0 or |b|<1e-14
 * (stresslets.cpp:174-178). */
void orc_dl_off(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sg, int nt, const double* tx,
                const double* tb, const double* tx0, const double* tn0,
                const double* tg0, double* out) {
  const double s8 = -6.0 * (1.0 / (8.0 * ORC_PI));
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double p[3] = {tx[3 * i], tx[3 * i + 1], tx[3 * i + 2]};
    double b = tb[i];
    double u[3] = {0, 0, 0};
    if (fabs(b) / DEL > 8.0) {
      sum_dl_plain(0, ns - 1, sx, sn, sa, sg, p, u);
      for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s8;
    } else {
      double c[3];
      orc_reg_coeff(b / DEL, c);
      sum_dl_reg(0, ns - 1, sx, sn, sa, sg, tg0 + 3 * i, tx0 + 3 * i, tn0 + 3 * i, b, c, DEL, p, u);
      for (int k = 0; k < 3; k++) {
        out[3 * i + k] = u[k] * s8;
        if (b < 0.0 || fabs(b) < 1.0e-14) out[3 * i + k] += tg0[3 * i + k];
      }
    }
  }
}

/* eval_Stokes_DL_onSurf + stokes_DL_7ord_onSurf (stresslets.cpp:186-279);
 * g0 = density of source i, jump +g0/2. */
void orc_dl_on(double DEL, int ns, const double* sx, const double* sn,
               const double* sa, const double* sg, int nt, const double* tx,
               double* out) {
  const double s8 = -6.0 * (1.0 / (8.0 * ORC_PI));
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double p[3] = {tx[3 * i], tx[3 * i + 1], tx[3 * i + 2]};
    const double* g0 = sg + 3 * i;
    double u[3] = {0, 0, 0};
    sum_dl_on(0, ns - 1, sx, sn, sa, sg, g0, DEL, p, u);
    for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s8 + 0.5 * g0[k];
  }
}

/* ======================================================================
 * Regularized pressure (highreg_stokes_extend_to_grid/stokes.cpp:99-203)
 * ====================================================================== */

/* Lapl_DL_7ord (extend_to_grid/utilities.cpp:190-211). */
void orc_lapl_dl_kernel(double r, double d, const double c[3], double* s2) {
  const double rp = rootpi();
  if (r < 1e-14) {
    *s2 = 256.0 / 15.0 / rp / (d * d * d);
  } else if (r / d > 8.0) {
    *s2 = 1.0 / (r * r * r);
  } else {
    double rd = r / d, rd2 = rd * rd, rd4 = rd2 * rd2;
    double ef = erf(rd);
    double rex = rd * exp(-rd2);
    double p2 = -2.0 / rp + 2.0 * (c[0] + 2.0 * (c[1] + c[2])) * rd2
                - 4.0 * (c[1] + 5.0 * c[2]) * rd4 + 8.0 * c[2] * rd4 * rd2;
    *s2 = (ef + p2 * rex) / (r * r * r);
  }
}

/* stokes_pressure_7ord (stokes.cpp:134-203): p = -1/(4 pi) sum_j
 * [-(f.n - f0.n0)(n.dx) - (n x f - n0 x f0).(n x dx)] H2 A_j, minus f0.n0
 * for b < -1e-12.  Every pair goes through Lapl_DL_7ord (no per-target far
 * class). */
void orc_pressure_off(double DEL, int ns, const double* sx, const double* sn,
                      const double* sa, const double* sf, int nt, const double* tx,
                      const double* tb, const double* tn0, const double* tf0, double* out) {
  const double s4 = 1.0 / (4.0 * ORC_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    const double* f0 = tf0 + 3 * i;
    const double* n0 = tn0 + 3 * i;
    double b = tb[i];
    double c[3];
    orc_reg_coeff(b / DEL, c);
    double f0n0 = f0[0] * n0[0] + f0[1] * n0[1] + f0[2] * n0[2];
    double m0[3] = {n0[1] * f0[2] - n0[2] * f0[1], n0[2] * f0[0] - n0[0] * f0[2],
                    n0[0] * f0[1] - n0[1] * f0[0]};
    double p = 0.0;
    for (int j = 0; j < ns; j++) {
      double dx[3] = {tx[3 * i] - sx[3 * j], tx[3 * i + 1] - sx[3 * j + 1], tx[3 * i + 2] - sx[3 * j + 2]};
      double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      double H2;
      orc_lapl_dl_kernel(r, DEL, c, &H2);
      const double* n = sn + 3 * j;
      const double* f = sf + 3 * j;
      double fn = f[0] * n[0] + f[1] * n[1] + f[2] * n[2];
      double nf[3] = {n[1] * f[2] - n[2] * f[1] - m0[0], n[2] * f[0] - n[0] * f[2] - m0[1],
                      n[0] * f[1] - n[1] * f[0] - m0[2]};
      double ndx = n[0] * dx[0] + n[1] * dx[1] + n[2] * dx[2];
      double ndx3[3] = {n[1] * dx[2] - n[2] * dx[1], n[2] * dx[0] - n[0] * dx[2], n[0] * dx[1] - n[1] * dx[0]};
      double normal = -(fn - f0n0) * ndx;
      double tangential = -(nf[0] * ndx3[0] + nf[1] * ndx3[1] + nf[2] * ndx3[2]);
      p += (normal + tangential) * H2 * sa[j];
    }
    p *= -s4;
    if (b < -1.e-12) p -= f0n0;
    out[i] = p;
  }
}

/* stokes_pressure_away (stokes.cpp:102-127): p = 1/(4 pi) sum_j A_j f.dx / r^3. */
void orc_pressure_away(int ns, const double* sx, const double* sa, const double* sf, int nt,
                       const double* tx, double* out) {
  const double s4 = 1.0 / (4.0 * ORC_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    double p = 0.0;
    for (int j = 0; j < ns; j++) {
      double dx = tx[3 * i] - sx[3 * j], dy = tx[3 * i + 1] - sx[3 * j + 1], dz = tx[3 * i + 2] - sx[3 * j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      double H2 = 1.0 / (r2 * sqrt(r2));
      H2 *= sf[3 * j] * dx + sf[3 * j + 1] * dy + sf[3 * j + 2] * dz;
      p += H2 * sa[j];
    }
    out[i] = p * s4;
  }
}

/* ======================================================================
 * KITC treecode
 * ====================================================================== */

typedef struct {
  long lo, hi;          /* members[0], members[1], inclusive (KITC.h:64) */
  double iv[6];         /* x0,x1,y0,y1,z0,z1 */
  double c[3];          /* xc, yc, zc */
  double mac;           /* (half diagonal)^2 / theta^2 */
  int child[8];
  int nchild;
  int parent;
} onode;

typedef struct {
  int P, N0, np1, pflat, ncomp;
  double sqth;
  int nn, cap;
  onode* nd;
  int n;
  double *x, *y, *z, *nx, *ny, *nz, *a, *l0, *l1, *l2;
  int* perm;
  double* t;    /* per node 3*(P+1): t1, t2, t3 */
  double* mom;  /* per node ncomp*pflat */
} otree;

static int tree_push(otree* T, const onode* n) {
  if (T->nn == T->cap) {
    T->cap = T->cap ? 2 * T->cap : 64;
    T->nd = (onode*)realloc(T->nd, sizeof(onode) * T->cap);
  }
  T->nd[T->nn] = *n;
  return T->nn++;
}

/* split_tree_node (KITC.cpp:116-322): eight midpoint children with the
 * reference's centre and MAC expressions; particles are partitioned stably by
 * octant (x fastest, "<= midpoint" goes low), which is the order produced by
 * the reference's address-table swaps (KITC.cpp:289-319); empty children are
 * dropped. */
static void tree_split(otree* T, int ni) {
  onode nd = T->nd[ni];
  double x0 = nd.iv[0], x1 = nd.iv[1], y0 = nd.iv[2], y1 = nd.iv[3], z0 = nd.iv[4], z1 = nd.iv[5];
  double mx = (x0 + x1) / 2.0, my = (y0 + y1) / 2.0, mz = (z0 + z1) / 2.0;
  double xc0 = (x0 + mx) / 2.0, xc1 = (x1 + mx) / 2.0;
  double yc0 = (y0 + my) / 2.0, yc1 = (y1 + my) / 2.0;
  double zc0 = (z0 + mz) / 2.0, zc1 = (z1 + mz) / 2.0;
  onode ch[8];
  for (int o = 0; o < 8; o++) {
    int hx = o & 1, hy = (o >> 1) & 1, hz = (o >> 2) & 1;
    memset(&ch[o], 0, sizeof(onode));
    ch[o].iv[0] = hx ? mx : x0; ch[o].iv[1] = hx ? x1 : mx;
    ch[o].iv[2] = hy ? my : y0; ch[o].iv[3] = hy ? y1 : my;
    ch[o].iv[4] = hz ? mz : z0; ch[o].iv[5] = hz ? z1 : mz;
    ch[o].c[0] = hx ? xc1 : xc0;
    ch[o].c[1] = hy ? yc1 : yc0;
    ch[o].c[2] = hz ? zc1 : zc0;
    double ex = hx ? (x1 - xc1) : (mx - xc0);
    double ey = hy ? (y1 - yc1) : (my - yc0);
    double ez = hz ? (z1 - zc1) : (mz - zc0);
    ch[o].mac = (ex * ex + ey * ey + ez * ez) / T->sqth;
    ch[o].parent = ni;
  }
  long lo = nd.lo, hi = nd.hi, n = hi - lo + 1;
  int* oct = (int*)malloc(sizeof(int) * n);
  long cnt[8] = {0};
  for (long k = 0; k < n; k++) {
    long p = lo + k;
    int o = (T->x[p] > mx ? 1 : 0) | (T->y[p] > my ? 2 : 0) | (T->z[p] > mz ? 4 : 0);
    oct[k] = o;
    cnt[o]++;
  }
  long start[8], s = lo;
  for (int o = 0; o < 8; o++) { start[o] = s; s += cnt[o]; }
  /* stable scatter of all per-particle arrays */
  double* arrs[10] = {T->x, T->y, T->z, T->nx, T->ny, T->nz, T->a, T->l0, T->l1, T->l2};
  double* tmp = (double*)malloc(sizeof(double) * n);
  int* itmp = (int*)malloc(sizeof(int) * n);
  for (int q = 0; q < 10; q++) {
    long pos[8];
    memcpy(pos, start, sizeof(pos));
    for (long k = 0; k < n; k++) tmp[pos[oct[k]]++ - lo] = arrs[q][lo + k];
    memcpy(arrs[q] + lo, tmp, sizeof(double) * n);
  }
  {
    long pos[8];
    memcpy(pos, start, sizeof(pos));
    for (long k = 0; k < n; k++) itmp[pos[oct[k]]++ - lo] = T->perm[lo + k];
    memcpy(T->perm + lo, itmp, sizeof(int) * n);
  }
  free(tmp);
  free(itmp);
  free(oct);
  int nc = 0;
  int kids[8];
  for (int o = 0; o < 8; o++) {
    if (cnt[o] >= 1) {
      ch[o].lo = start[o];
      ch[o].hi = start[o] + cnt[o] - 1;
      kids[nc++] = tree_push(T, &ch[o]);
    }
  }
  T->nd[ni].nchild = nc;
  for (int k = 0; k < nc; k++) T->nd[ni].child[k] = kids[k];
}

/* build_tree_3D_Recursive (KITC.cpp:325-348): split iff n >= N0. */
static void tree_build_rec(otree* T, int ni) {
  long n = T->nd[ni].hi - T->nd[ni].lo + 1;
  if (n >= (long)T->N0) {
    tree_split(T, ni);
    int nc = T->nd[ni].nchild;
    int kids[8];
    memcpy(kids, T->nd[ni].child, sizeof(kids));
    for (int k = 0; k < nc; k++) tree_build_rec(T, kids[k]);
  }
}

/* Cluster_Chev_Points (KITC.cpp:780-804). */
static void tree_cheb(otree* T) {
  int P = T->P;
  double hh = 3.14159265358979323846 / P;
  double* t = (double*)malloc(sizeof(double) * (P + 1));
  for (int i = 0; i <= P; i++) t[i] = cos(i * hh);
  T->t = (double*)malloc(sizeof(double) * 3 * (P + 1) * T->nn);
  for (int q = 0; q < T->nn; q++) {
    double* tq = T->t + 3 * (P + 1) * q;
    const double* iv = T->nd[q].iv;
    for (int i = 0; i <= P; i++) {
      tq[i] = iv[0] + (t[i] + 1) / 2 * (iv[1] - iv[0]);
      tq[(P + 1) + i] = iv[2] + (t[i] + 1) / 2 * (iv[3] - iv[2]);
      tq[2 * (P + 1) + i] = iv[4] + (t[i] + 1) / 2 * (iv[5] - iv[4]);
    }
  }
  free(t);
}

/* Panel_Moments_SL / Panel_Moments_DL (KITC.cpp:351-668): barycentric weights
 * w_j = (-1)^j delta_j (delta_0 = delta_P = 1/2), a_j = w_j/(y - t_j) with a
 * snap to the node when |y - t_j| < 1e-7, D = 1/(sum a1 sum a2 sum a3), and
 * W[c][kk] = sum_s a1_i a2_j a3_k D A_s q_{s,c}. */
static void tree_moments_node(otree* T, int q) {
  int P = T->P, np1 = T->np1, pflat = T->pflat, ncomp = T->ncomp;
  long lo = T->nd[q].lo, hi = T->nd[q].hi, n = hi - lo + 1;
  const double* t1 = T->t + 3 * np1 * q;
  const double* t2 = t1 + np1;
  const double* t3 = t2 + np1;
  double* w = (double*)malloc(sizeof(double) * np1);
  for (int j = 0; j <= P; j++) {
    double dj = (j == 0 || j == P) ? 0.5 : 1.0;
    w[j] = ((j % 2 == 0) ? 1 : -1) * dj;
  }
  double* A1 = (double*)calloc((size_t)np1 * n, sizeof(double));
  double* A2 = (double*)calloc((size_t)np1 * n, sizeof(double));
  double* A3 = (double*)calloc((size_t)np1 * n, sizeof(double));
  double* D = (double*)calloc(n, sizeof(double));
  for (long s = 0; s < n; s++) {
    double x = T->x[lo + s], y = T->y[lo + s], z = T->z[lo + s];
    int f1 = -1, f2 = -1, f3 = -1;
    double S1 = 0, S2 = 0, S3 = 0;
    for (int j = 0; j <= P; j++) {
      if (fabs(x - t1[j]) < 1.e-07) f1 = j;
      else { double v = w[j] / (x - t1[j]); A1[j * n + s] = v; S1 += v; }
      if (fabs(y - t2[j]) < 1.e-07) f2 = j;
      else { double v = w[j] / (y - t2[j]); A2[j * n + s] = v; S2 += v; }
      if (fabs(z - t3[j]) < 1.e-07) f3 = j;
      else { double v = w[j] / (z - t3[j]); A3[j * n + s] = v; S3 += v; }
    }
    if (f1 > -1) { S1 = 1.0; for (int j = 0; j <= P; j++) A1[j * n + s] = 0.0; A1[f1 * n + s] = 1.0; }
    if (f2 > -1) { S2 = 1.0; for (int j = 0; j <= P; j++) A2[j * n + s] = 0.0; A2[f2 * n + s] = 1.0; }
    if (f3 > -1) { S3 = 1.0; for (int j = 0; j <= P; j++) A3[j * n + s] = 0.0; A3[f3 * n + s] = 1.0; }
    D[s] = 1.0 / (S1 * S2 * S3);
  }
  double* m = T->mom + (size_t)ncomp * pflat * q;
  int kk = -1;
  for (int i = 0; i <= P; i++)
    for (int j = 0; j <= P; j++)
      for (int k = 0; k <= P; k++) {
        kk++;
        double sum[12] = {0};
        for (long s = 0; s < n; s++) {
          long p = lo + s;
          double sv = A1[i * n + s] * A2[j * n + s] * A3[k * n + s] * D[s];
          sv *= T->a[p];
          if (ncomp == 6) {
            sum[0] += sv * T->l0[p];
            sum[1] += sv * T->l1[p];
            sum[2] += sv * T->l2[p];
            sum[3] += sv * T->nx[p];
            sum[4] += sv * T->ny[p];
            sum[5] += sv * T->nz[p];
          } else {
            double l[3] = {T->l0[p], T->l1[p], T->l2[p]};
            for (int c = 0; c < 3; c++) {
              sum[3 * c] += sv * l[c] * T->nx[p];
              sum[3 * c + 1] += sv * l[c] * T->ny[p];
              sum[3 * c + 2] += sv * l[c] * T->nz[p];
            }
            sum[9] += sv * T->nx[p];
            sum[10] += sv * T->ny[p];
            sum[11] += sv * T->nz[p];
          }
        }
        for (int c = 0; c < ncomp; c++) m[(size_t)c * pflat + kk] = sum[c];
      }
  free(A1); free(A2); free(A3); free(D); free(w);
}

/* Copies the inputs into tree-owned arrays and builds the tree
 * (eval_Stokes_*_tree drivers, KITC.cpp:1409-1442). */
static void tree_create(otree* T, const orc_tree_params* tp, int dl, int ns,
                        const double* sx, const double* sn, const double* sa,
                        const double* sd) {
  memset(T, 0, sizeof(*T));
  T->P = tp->P;
  T->N0 = tp->N0;
  T->np1 = tp->P + 1;
  T->pflat = T->np1 * T->np1 * T->np1;
  T->ncomp = dl ? 12 : 6;
  T->sqth = tp->theta * tp->theta;
  T->n = ns;
  double** arrs[10] = {&T->x, &T->y, &T->z, &T->nx, &T->ny, &T->nz, &T->a, &T->l0, &T->l1, &T->l2};
  for (int q = 0; q < 10; q++) *arrs[q] = (double*)malloc(sizeof(double) * (ns > 0 ? ns : 1));
  T->perm = (int*)malloc(sizeof(int) * (ns > 0 ? ns : 1));
  for (int i = 0; i < ns; i++) {
    T->x[i] = sx[3 * i]; T->y[i] = sx[3 * i + 1]; T->z[i] = sx[3 * i + 2];
    if (sn) { T->nx[i] = sn[3 * i]; T->ny[i] = sn[3 * i + 1]; T->nz[i] = sn[3 * i + 2]; }
    else { T->nx[i] = T->ny[i] = T->nz[i] = 0; }
    T->a[i] = sa ? sa[i] : 0;
    if (sd) { T->l0[i] = sd[3 * i]; T->l1[i] = sd[3 * i + 1]; T->l2[i] = sd[3 * i + 2]; }
    else { T->l0[i] = T->l1[i] = T->l2[i] = 0; }
    T->perm[i] = i;
  }
  /* build_tree_init (KITC.cpp:30-69) */
  onode root;
  memset(&root, 0, sizeof(root));
  root.lo = 0;
  root.hi = (long)ns - 1;
  double eps = 1.e-8;
  for (int d = 0; d < 3; d++) {
    root.iv[2 * d] = -1.0 - eps;
    root.iv[2 * d + 1] = 1.0 + eps;
    root.iv[2 * d] += tp->ctr[d];
    root.iv[2 * d + 1] += tp->ctr[d];
    root.c[d] = 0.5 * (root.iv[2 * d] + root.iv[2 * d + 1]);
  }
  root.mac = (3 * 10 * 10 / 4) / T->sqth;
  root.parent = -1;
  tree_push(T, &root);
  tree_build_rec(T, 0);
  tree_cheb(T);
  T->mom = (double*)calloc((size_t)T->ncomp * T->pflat * T->nn, sizeof(double));
}

static void tree_compute_moments(otree* T) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 1; q < T->nn; q++) tree_moments_node(T, q);  /* root skipped (KITC.cpp:1446) */
}

static void tree_free(otree* T) {
  double* arrs[10] = {T->x, T->y, T->z, T->nx, T->ny, T->nz, T->a, T->l0, T->l1, T->l2};
  for (int q = 0; q < 10; q++) free(arrs[q]);
  free(T->perm);
  free(T->nd);
  free(T->t);
  free(T->mom);
}

/* Call_Treecode_SL (KITC.cpp:671-721): plain Stokeslet from the proxies with
 * target-side subtraction W_f - (f0.n0) W_n. */
static void proxy_sl(const otree* T, int q, const double p[3], double f0n0, double u[3]) {
  int np1 = T->np1, pflat = T->pflat;
  const double* t1 = T->t + 3 * np1 * q;
  const double* m = T->mom + (size_t)T->ncomp * pflat * q;
  double dxa[16], dya[16], dza[16];
  for (int i = 0; i < np1; i++) {
    dxa[i] = p[0] - t1[i];
    dya[i] = p[1] - t1[np1 + i];
    dza[i] = p[2] - t1[2 * np1 + i];
  }
  int kk = -1;
  for (int i = 0; i < np1; i++) {
    double d0 = dxa[i], ti = d0 * d0;
    for (int j = 0; j < np1; j++) {
      double d1 = dya[j], tj = d1 * d1;
      for (int k = 0; k < np1; k++) {
        double d2 = dza[k], tk = d2 * d2;
        kk++;
        double m0 = m[kk] - f0n0 * m[3 * pflat + kk];
        double m1 = m[pflat + kk] - f0n0 * m[4 * pflat + kk];
        double m2 = m[2 * pflat + kk] - f0n0 * m[5 * pflat + kk];
        double r2 = ti + tj + tk;
        double r = sqrt(r2);
        double tmp = (m0 * d0 + m1 * d1 + m2 * d2) / (r2 * r);
        u[0] += m0 / r + tmp * d0;
        u[1] += m1 / r + tmp * d1;
        u[2] += m2 / r + tmp * d2;
      }
    }
  }
}

/* Call_Treecode_DL (KITC.cpp:724-777): plain stresslet from the proxies with
 * target-side subtraction W_{g_l n} - g0_l W_n. */
static void proxy_dl(const otree* T, int q, const double p[3], const double g0[3], double u[3]) {
  int np1 = T->np1, pflat = T->pflat;
  const double* t1 = T->t + 3 * np1 * q;
  const double* m = T->mom + (size_t)T->ncomp * pflat * q;
  double dxa[16], dya[16], dza[16];
  for (int i = 0; i < np1; i++) {
    dxa[i] = p[0] - t1[i];
    dya[i] = p[1] - t1[np1 + i];
    dza[i] = p[2] - t1[2 * np1 + i];
  }
  int kk = -1;
  for (int i = 0; i < np1; i++) {
    double d0 = dxa[i], ti = d0 * d0;
    for (int j = 0; j < np1; j++) {
      double d1 = dya[j], tj = d1 * d1;
      for (int k = 0; k < np1; k++) {
        double d2 = dza[k], tk = d2 * d2;
        kk++;
        double tmp[3];
        for (int l = 0; l < 3; l++) {
          double a1 = m[(size_t)(3 * l) * pflat + kk] - g0[l] * m[(size_t)9 * pflat + kk];
          double a2 = m[(size_t)(3 * l + 1) * pflat + kk] - g0[l] * m[(size_t)10 * pflat + kk];
          double a3 = m[(size_t)(3 * l + 2) * pflat + kk] - g0[l] * m[(size_t)11 * pflat + kk];
          tmp[l] = a1 * d0 + a2 * d1 + a3 * d2;
        }
        double r2 = ti + tj + tk;
        double r = sqrt(r2);
        double kern = (tmp[0] * d0 + tmp[1] * d1 + tmp[2] * d2) / (r2 * r2 * r);
        u[0] += d0 * kern;
        u[1] += d1 * kern;
        u[2] += d2 * kern;
      }
    }
  }
}

enum { K_SL_ON = 0, K_SL_OFF = 1, K_DL_ON = 2, K_DL_OFF = 3 };

typedef struct {
  int kind;
  double DEL;
  double p[3];
  double b;
  double f0n0;
  double g0[3], x0[3], n0[3];
  double c[3];
} otarget;

/* Leaf direct sums Call_Ds_SL_onSurf / Call_Ds_SL / Call_Ds_DL_onSurf /
 * Call_Ds_DL (KITC.cpp:807-1062).  The off-surface ones use the plain kernel
 * without subtraction when |b|/delta > 8 (KITC.cpp:856, 970). */
static void leaf_sum(const otree* T, long lo, long hi, const otarget* g, double u[3]) {
  /* tree arrays are SoA; build interleaved views on the fly */
  for (long j = lo; j <= hi; j++) {
    double sx[3] = {T->x[j], T->y[j], T->z[j]};
    double sn[3] = {T->nx[j], T->ny[j], T->nz[j]};
    double sd[3] = {T->l0[j], T->l1[j], T->l2[j]};
    double sa = T->a[j];
    switch (g->kind) {
      case K_SL_ON:
        sum_sl_reg(1, 0, 0, sx, sn, &sa, sd, g->f0n0, NULL, g->DEL, g->p, u);
        break;
      case K_SL_OFF:
        if (fabs(g->b) / g->DEL > 8.0) sum_sl_plain(0, 0, sx, &sa, sd, g->p, u);
        else sum_sl_reg(0, 0, 0, sx, sn, &sa, sd, g->f0n0, g->c, g->DEL, g->p, u);
        break;
      case K_DL_ON:
        sum_dl_on(0, 0, sx, sn, &sa, sd, g->g0, g->DEL, g->p, u);
        break;
      default:
        if (fabs(g->b) / g->DEL > 8.0) sum_dl_plain(0, 0, sx, sn, &sa, sd, g->p, u);
        else {
          /* Call_Ds_DL accumulates m1*H2 + m2*H3 per component in one add
           * (KITC.cpp:1043-1057); the regularized helper adds the two parts
           * separately, which only reorders a rounding. */
          sum_dl_reg(0, 0, sx, sn, &sa, sd, g->g0, g->x0, g->n0, g->b, g->c, g->DEL, g->p, u);
        }
        break;
    }
  }
}

/* eval_{SL,DL}_Sum_cluster{,_onSurf} (KITC.cpp:1066-1386): far if MAC < R^2,
 * leaf if (members[1] - members[0]) < N0, else recurse into the children in
 * creation order.  Each level sums its children's results. */
static void traverse(const otree* T, int q, const otarget* g, double u[3]) {
  const onode* nd = &T->nd[q];
  double tpx = g->p[0] - nd->c[0], tpy = g->p[1] - nd->c[1], tpz = g->p[2] - nd->c[2];
  double R2 = tpx * tpx + tpy * tpy + tpz * tpz;
  if (nd->mac < R2) {
    double v[3] = {0, 0, 0};
    if (g->kind == K_SL_ON || g->kind == K_SL_OFF) proxy_sl(T, q, g->p, g->f0n0, v);
    else proxy_dl(T, q, g->p, g->g0, v);
    for (int k = 0; k < 3; k++) u[k] += v[k];
  } else if ((nd->hi - nd->lo) < (long)T->N0) {
    double v[3] = {0, 0, 0};
    leaf_sum(T, nd->lo, nd->hi, g, v);
    for (int k = 0; k < 3; k++) u[k] += v[k];
  } else {
    double s[3] = {0, 0, 0};
    for (int c = 0; c < nd->nchild; c++) {
      double v[3] = {0, 0, 0};
      traverse(T, nd->child[c], g, v);
      for (int k = 0; k < 3; k++) s[k] += v[k];
    }
    for (int k = 0; k < 3; k++) u[k] += s[k];
  }
}

static void tree_eval(int kind, const orc_tree_params* tp, double DEL, int ns,
                      const double* sx, const double* sn, const double* sa,
                      const double* sd, int nt, const double* tx, const double* tb,
                      const double* tx0, const double* tn0, const double* td0,
                      double* out) {
  otree T;
  int dl = (kind == K_DL_ON || kind == K_DL_OFF);
  tree_create(&T, tp, dl, ns, sx, sn, sa, sd);
  tree_compute_moments(&T);
  const double s8 = 1.0 / (8.0 * ORC_PI);
  const double s86 = -6.0 / (8.0 * ORC_PI);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    otarget g;
    memset(&g, 0, sizeof(g));
    g.kind = kind;
    g.DEL = DEL;
    for (int k = 0; k < 3; k++) g.p[k] = tx[3 * i + k];
    if (kind == K_SL_ON) {
      g.f0n0 = sd[3 * i] * sn[3 * i] + sd[3 * i + 1] * sn[3 * i + 1] + sd[3 * i + 2] * sn[3 * i + 2];
    } else if (kind == K_SL_OFF) {
      g.b = tb[i];
      g.f0n0 = td0[3 * i] * tn0[3 * i] + td0[3 * i + 1] * tn0[3 * i + 1] + td0[3 * i + 2] * tn0[3 * i + 2];
      if (!(fabs(g.b) / DEL > 8.0)) orc_reg_coeff(g.b / DEL, g.c);
    } else if (kind == K_DL_ON) {
      for (int k = 0; k < 3; k++) g.g0[k] = sd[3 * i + k];
    } else {
      g.b = tb[i];
      for (int k = 0; k < 3; k++) {
        g.g0[k] = td0[3 * i + k];
        g.x0[k] = tx0[3 * i + k];
        g.n0[k] = tn0[3 * i + k];
      }
      if (!(fabs(g.b) / DEL > 8.0)) orc_reg_coeff(g.b / DEL, g.c);
    }
    double u[3] = {0, 0, 0};
    traverse(&T, 0, &g, u);
    if (kind == K_SL_ON || kind == K_SL_OFF) {
      for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s8;
    } else if (kind == K_DL_ON) {
      for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s86 + 0.5 * g.g0[k];  /* KITC.cpp:1693 */
    } else {
      double xi = 0.0;                         /* KITC.cpp:1805-1811 */
      if (fabs(g.b) < 1.0e-12) xi = 0.5;
      else if (g.b < 0.0) xi = 1.0;
      for (int k = 0; k < 3; k++) out[3 * i + k] = u[k] * s86 + xi * g.g0[k];
    }
  }
  tree_free(&T);
}

void orc_sl_on_tree(const orc_tree_params* tp, double DEL, int ns, const double* sx,
                    const double* sn, const double* sa, const double* sf, int nt,
                    const double* tx, double* out) {
  tree_eval(K_SL_ON, tp, DEL, ns, sx, sn, sa, sf, nt, tx, NULL, NULL, NULL, NULL, out);
}
void orc_sl_off_tree(const orc_tree_params* tp, double DEL, int ns, const double* sx,
                     const double* sn, const double* sa, const double* sf, int nt,
                     const double* tx, const double* tb, const double* tn0,
                     const double* tf0, double* out) {
  tree_eval(K_SL_OFF, tp, DEL, ns, sx, sn, sa, sf, nt, tx, tb, NULL, tn0, tf0, out);
}
void orc_dl_on_tree(const orc_tree_params* tp, double DEL, int ns, const double* sx,
                    const double* sn, const double* sa, const double* sg, int nt,
                    const double* tx, double* out) {
  tree_eval(K_DL_ON, tp, DEL, ns, sx, sn, sa, sg, nt, tx, NULL, NULL, NULL, NULL, out);
}
void orc_dl_off_tree(const orc_tree_params* tp, double DEL, int ns, const double* sx,
                     const double* sn, const double* sa, const double* sg, int nt,
                     const double* tx, const double* tb, const double* tx0,
                     const double* tn0, const double* tg0, double* out) {
  tree_eval(K_DL_OFF, tp, DEL, ns, sx, sn, sa, sg, nt, tx, tb, tx0, tn0, tg0, out);
}

int orc_tree_structure(const orc_tree_params* tp, int ns, const double* sx, int cap,
                       int* members, double* interval, double* centre, double* mac,
                       int* parent, int* perm) {
  otree T;
  tree_create(&T, tp, 0, ns, sx, NULL, NULL, NULL);
  int nn = T.nn;
  if (cap >= nn) {
    for (int q = 0; q < nn; q++) {
      members[2 * q] = (int)T.nd[q].lo;
      members[2 * q + 1] = (int)T.nd[q].hi;
      memcpy(interval + 6 * q, T.nd[q].iv, 6 * sizeof(double));
      memcpy(centre + 3 * q, T.nd[q].c, 3 * sizeof(double));
      mac[q] = T.nd[q].mac;
      parent[q] = T.nd[q].parent;
    }
    memcpy(perm, T.perm, sizeof(int) * ns);
  }
  tree_free(&T);
  return nn;
}

int orc_tree_moments(const orc_tree_params* tp, int dl, int ns, const double* sx,
                     const double* sn, const double* sa, const double* sdens,
                     int cap_nodes, double* moments) {
  otree T;
  tree_create(&T, tp, dl, ns, sx, sn, sa, sdens);
  int nn = T.nn;
  if (cap_nodes >= nn) {
    tree_compute_moments(&T);
    memcpy(moments, T.mom, sizeof(double) * (size_t)T.ncomp * T.pflat * nn);
  }
  tree_free(&T);
  return nn;
}

/* ======================================================================
 * Stencil and interpolation (highreg_stokes_2surf/main.cpp:682-778, 916-1046)
 * ====================================================================== */

/* sixPoint_stencil (main.cpp:682-748) for one target: scans every source point
 * (the Monge-range restriction is overridden at main.cpp:969-970) and keeps
 * the last match per slot; unmatched slots stay 0 (the reference's "unset"). */
static void stencil_one(double h, int ns, const double* sx, int ind1, int ind2,
                        int ind3, double z0, double zctr, double xLow, double yLow,
                        double xHigh, double yHigh, int st[36]) {
  const double tol = 1.e-14;
  double cx[6] = {xLow - 2.0 * h, xLow - h, xLow, xHigh, xHigh + h, xHigh + 2.0 * h};
  double cy[6] = {yLow - 2.0 * h, yLow - h, yLow, yHigh, yHigh + h, yHigh + 2.0 * h};
  for (int k = 0; k < 36; k++) st[k] = 0;
  for (int j = 0; j < ns; j++) {
    double z = sx[3 * j + ind3] - zctr;
    if (!(((z0 > 0) && (z > 0)) || ((z0 < 0) && (z < 0)))) continue;
    double X = sx[3 * j + ind1], Y = sx[3 * j + ind2];
    int col = -1;
    for (int c = 0; c < 6; c++)
      if (fabs(cx[c] - X) < tol) { col = c; break; }
    if (col < 0) continue;
    for (int r = 0; r < 6; r++)
      if (fabs(cy[r] - Y) < tol) { st[col * 6 + r] = j; break; }
  }
}

/* sixPoint_coeff (main.cpp:750-757). */
static void six_coeff(double p, double a[6]) {
  a[0] = -p * (p * p - 1.0) * (p - 2.0) * (p - 3.0) / 120.0;
  a[1] = p * (p - 1.0) * (p * p - 4.0) * (p - 3.0) / 24.0;
  a[2] = -(p * p - 1.0) * (p * p - 4.0) * (p - 3.0) / 12.0;
  a[3] = p * (p + 1.0) * (p * p - 4.0) * (p - 3.0) / 12.0;
  a[4] = -p * (p * p - 1.0) * (p + 2.0) * (p - 3.0) / 24.0;
  a[5] = p * (p * p - 1.0) * (p * p - 4.0) / 120.0;
}

static void target_cell(const double ctr[3], double h, const double* x0, const int* i123,
                        double* xLow, double* yLow, double* z0) {
  int ind1 = i123[0], ind2 = i123[1], ind3 = i123[2];
  *xLow = floor((x0[ind1] - ctr[ind1]) / h) * h + ctr[ind1];
  *yLow = floor((x0[ind2] - ctr[ind2]) / h) * h + ctr[ind2];
  *z0 = x0[ind3] - ctr[ind3];
}

void orc_stencil(const double ctr[3], double h, int ns, const double* sx, int nt,
                 const int* nrst, const double* tx0, const int* i123, int* stencil) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    int* st = stencil + 36 * i;
    if (nrst[i] == 0) {
      for (int k = 0; k < 36; k++) st[k] = 0;
      continue;
    }
    double xLow, yLow, z0;
    target_cell(ctr, h, tx0 + 3 * i, i123 + 3 * i, &xLow, &yLow, &z0);
    stencil_one(h, ns, sx, i123[3 * i], i123[3 * i + 1], i123[3 * i + 2], z0,
                ctr[i123[3 * i + 2]], xLow, yLow, xLow + h, yLow + h, st);
  }
}

/* interpolate_DL_density (main.cpp:916-1046) with sixPointInterpolate
 * (main.cpp:759-778): g0 = sum_j b_j sum_i a_i g[stencil[i][j]]. */
void orc_interp_g0(const double ctr[3], double h, int ns, const double* sx,
                   const double* sg, int nt, const int* nrst, const double* tx0,
                   const int* i123, double* g0) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; i++) {
    g0[3 * i] = g0[3 * i + 1] = g0[3 * i + 2] = 0;
    if (nrst[i] == 0) continue;
    double xLow, yLow, z0;
    const int* ii = i123 + 3 * i;
    target_cell(ctr, h, tx0 + 3 * i, ii, &xLow, &yLow, &z0);
    int st[36];
    stencil_one(h, ns, sx, ii[0], ii[1], ii[2], z0, ctr[ii[2]], xLow, yLow, xLow + h,
                yLow + h, st);
    double p = (tx0[3 * i + ii[0]] - xLow) / h;
    double q = (tx0[3 * i + ii[1]] - yLow) / h;
    double a[6], bb[6];
    six_coeff(p, a);
    six_coeff(q, bb);
    for (int c = 0; c < 3; c++) {
      double fj[6];
      for (int j = 0; j < 6; j++) {
        fj[j] = 0.0;
        for (int k = 0; k < 6; k++) fj[j] += a[k] * sg[3 * st[k * 6 + j] + c];
      }
      double v = 0.0;
      for (int j = 0; j < 6; j++) v += bb[j] * fj[j];
      g0[3 * i + c] = v;
    }
  }
}

/* ======================================================================
 * Two-surface problem: matvec, RHS, GMRES (main.cpp:140-677)
 * ====================================================================== */

static void op_dl_on(const orc_problem* pb, int s, const double* g, double* out) {
  const orc_surface* S = &pb->s[s];
  if (pb->use_tree)
    orc_dl_on_tree(&pb->tree[s], pb->DEL, S->n, S->x, S->nrml, S->area, g, S->n, S->x, out);
  else
    orc_dl_on(pb->DEL, S->n, S->x, S->nrml, S->area, g, S->n, S->x, out);
}

/* DL of surface s (density g) at the points of the other surface t, whose
 * nearest data is relative to s and whose g0 was interpolated from g. */
static void op_dl_off(const orc_problem* pb, int s, const double* g, const double* g0, double* out) {
  const orc_surface* S = &pb->s[s];
  const orc_surface* T = &pb->s[1 - s];
  if (pb->use_tree)
    orc_dl_off_tree(&pb->tree[s], pb->DEL, S->n, S->x, S->nrml, S->area, g, T->n, T->x,
                    T->b, T->x0, T->n0, g0, out);
  else
    orc_dl_off(pb->DEL, S->n, S->x, S->nrml, S->area, g, T->n, T->x, T->b, T->x0, T->n0, g0, out);
}

void orc_matvec(const orc_problem* pb, const double* u, double* Au) {
  const double factor = 2.0 / 3.0;
  int n1 = pb->s[0].n, n2 = pb->s[1].n;
  const double* g1 = u;            /* main.cpp:554-563 */
  const double* g2 = u + 3 * n1;
  double* g0_t2 = (double*)malloc(sizeof(double) * 3 * n2);  /* surface-1 density at targets 2 */
  double* g0_t1 = (double*)malloc(sizeof(double) * 3 * n1);
  const orc_surface *S1 = &pb->s[0], *S2 = &pb->s[1];
  orc_interp_g0(S1->ctr, pb->h, n1, S1->x, g1, n2, S2->nrst, S2->x0, S2->i123, g0_t2);
  orc_interp_g0(S2->ctr, pb->h, n2, S2->x, g2, n1, S1->nrst, S1->x0, S1->i123, g0_t1);
  double* on1 = (double*)malloc(sizeof(double) * 3 * n1);
  double* c1 = (double*)malloc(sizeof(double) * 3 * n2);
  double* c2 = (double*)malloc(sizeof(double) * 3 * n1);
  double* on2 = (double*)malloc(sizeof(double) * 3 * n2);
  op_dl_on(pb, 0, g1, on1);
  op_dl_off(pb, 0, g1, g0_t2, c1);
  for (int i = 0; i < 3 * n1; i++) Au[i] = u[i] - on1[i] * factor;            /* main.cpp:612-616 */
  for (int i = 0; i < 3 * n2; i++) Au[3 * n1 + i] = u[3 * n1 + i] - c1[i] * factor;
  op_dl_off(pb, 1, g2, g0_t1, c2);
  op_dl_on(pb, 1, g2, on2);
  for (int i = 0; i < 3 * n1; i++) Au[i] -= c2[i] * factor;                   /* main.cpp:667-676 */
  for (int i = 0; i < 3 * n2; i++) Au[3 * n1 + i] -= on2[i] * factor;
  free(g0_t2); free(g0_t1); free(on1); free(c1); free(c2); free(on2);
}

void orc_eval_sl(const orc_problem* pb, double* sl) {
  const orc_surface *S1 = &pb->s[0], *S2 = &pb->s[1];
  int n1 = S1->n, n2 = S2->n;
  double* on1 = (double*)malloc(sizeof(double) * 3 * n1);
  double* c1 = (double*)malloc(sizeof(double) * 3 * n2);
  double* c2 = (double*)malloc(sizeof(double) * 3 * n1);
  double* on2 = (double*)malloc(sizeof(double) * 3 * n2);
  if (pb->use_tree) {
    orc_sl_on_tree(&pb->tree[0], pb->DEL, n1, S1->x, S1->nrml, S1->area, S1->f, n1, S1->x, on1);
    orc_sl_off_tree(&pb->tree[0], pb->DEL, n1, S1->x, S1->nrml, S1->area, S1->f, n2, S2->x, S2->b, S2->n0, S2->f0, c1);
    orc_sl_off_tree(&pb->tree[1], pb->DEL, n2, S2->x, S2->nrml, S2->area, S2->f, n1, S1->x, S1->b, S1->n0, S1->f0, c2);
    orc_sl_on_tree(&pb->tree[1], pb->DEL, n2, S2->x, S2->nrml, S2->area, S2->f, n2, S2->x, on2);
  } else {
    orc_sl_on(pb->DEL, n1, S1->x, S1->nrml, S1->area, S1->f, n1, S1->x, on1);
    orc_sl_off(pb->DEL, n1, S1->x, S1->nrml, S1->area, S1->f, n2, S2->x, S2->b, S2->n0, S2->f0, c1);
    orc_sl_off(pb->DEL, n2, S2->x, S2->nrml, S2->area, S2->f, n1, S1->x, S1->b, S1->n0, S1->f0, c2);
    orc_sl_on(pb->DEL, n2, S2->x, S2->nrml, S2->area, S2->f, n2, S2->x, on2);
  }
  for (int i = 0; i < 3 * n1; i++) { sl[i] = on1[i]; sl[i] += c2[i]; }      /* main.cpp:478-536 */
  for (int i = 0; i < 3 * n2; i++) { sl[3 * n1 + i] = c1[i]; sl[3 * n1 + i] += on2[i]; }
  free(on1); free(c1); free(c2); free(on2);
}

/* GMRES(m) as in main.cpp:210-416: MGS Arnoldi, Givens rotations, residual
 * estimate |g_{j+1}|/|b|, x rebuilt after every inner step. */
int orc_gmres(const orc_problem* pb, int n, const double* b, double* x, int im,
              double tol, int iter_max, int fix_restart, double* resid_hist,
              double* final_resid) {
  double* V = (double*)malloc(sizeof(double) * (size_t)n * (im + 1));
  double* H = (double*)calloc((size_t)(im + 1) * im, sizeof(double));
  double* g = (double*)calloc(im + 1, sizeof(double));
  double* y = (double*)calloc(im + 1, sizeof(double));
  double* c = (double*)calloc(im, sizeof(double));
  double* s = (double*)calloc(im, sizeof(double));
  double* Ax = (double*)calloc(n, sizeof(double));
  double* xbase = (double*)calloc(n, sizeof(double));
  double* w = (double*)malloc(sizeof(double) * n);
#define HH(i, j) H[(size_t)(i) * im + (j)]
#define VV(k, i) V[(size_t)(i) * n + (k)]
  for (int i = 0; i < n; i++) x[i] = 0.0;
  int iter = 0;
  double error = 1.0;
  while (iter <= iter_max) {
    if (error <= tol || iter == iter_max) break;
    if (iter > 0) orc_matvec(pb, x, Ax);
    for (int i = 0; i < n; i++) VV(i, 0) = fix_restart ? (b[i] - Ax[i]) : b[i];  /* main.cpp:286 */
    if (fix_restart) memcpy(xbase, x, sizeof(double) * n);
    double r_norm = 0.0, b_norm = 0.0;
    for (int i = 0; i < n; i++) { r_norm += VV(i, 0) * VV(i, 0); b_norm += b[i] * b[i]; }
    r_norm = sqrt(r_norm);
    b_norm = sqrt(b_norm);
    error = r_norm;
    if (b_norm > 1.e-8) error /= b_norm;
    double tmp = 1.0 / r_norm;
    for (int i = 0; i < n; i++) VV(i, 0) = VV(i, 0) * tmp;
    g[0] = r_norm;
    int j = -1;
    while ((j < im - 1) && (error > tol) && (iter < iter_max)) {
      j = j + 1;
      iter = iter + 1;
      orc_matvec(pb, &VV(0, j), w);
      for (int i = 0; i < n; i++) VV(i, j + 1) = w[i];
      for (int i = 0; i <= j; i++) {                                      /* MGS main.cpp:340-345 */
        double t = 0.0;
        for (int k = 0; k < n; k++) t += VV(k, i) * VV(k, j + 1);
        HH(i, j) = t;
        for (int k = 0; k < n; k++) VV(k, j + 1) -= t * VV(k, i);
      }
      double vn = 0.0;
      for (int i = 0; i < n; i++) vn += VV(i, j + 1) * VV(i, j + 1);
      vn = sqrt(vn);
      HH(j + 1, j) = vn;
      if (vn != 0.0) {
        vn = 1.0 / vn;
        for (int k = 0; k < n; k++) VV(k, j + 1) *= vn;
      }
      if (j != 0) {                                                       /* main.cpp:362-368 */
        for (int i = 0; i <= j - 1; i++) {
          double t = HH(i, j);
          HH(i, j) = c[i] * t + s[i] * HH(i + 1, j);
          HH(i + 1, j) = -s[i] * t + c[i] * HH(i + 1, j);
        }
      }
      double hh = sqrt(HH(j, j) * HH(j, j) + HH(j + 1, j) * HH(j + 1, j));
      if (hh == 0.0) hh = 1.e-16;
      c[j] = HH(j, j) / hh;
      s[j] = HH(j + 1, j) / hh;
      HH(j, j) = c[j] * HH(j, j) + s[j] * HH(j + 1, j);
      for (int i = 0; i <= j; i++) HH(j + 1, i) = 0.0;
      g[j + 1] = -s[j] * g[j];
      g[j] = c[j] * g[j];
      error = fabs(g[j + 1]);
      if (b_norm > 1.e-8) error /= b_norm;
      if (resid_hist) resid_hist[iter - 1] = error;
      for (int i = 0; i <= j; i++) y[i] = g[i];                          /* main.cpp:402-414 */
      y[j] = y[j] / HH(j, j);
      for (int k = j - 1; k >= 0; k--) {
        double t = y[k];
        for (int i = k + 1; i <= j; i++) t -= HH(k, i) * y[i];
        y[k] = t / HH(k, k);
      }
      for (int i = 0; i < n; i++) x[i] = fix_restart ? xbase[i] : 0.0;
      for (int i = 0; i <= j; i++)
        for (int k = 0; k < n; k++) x[k] += y[i] * VV(k, i);
    }
  }
#undef HH
#undef VV
  if (final_resid) *final_resid = error;
  free(V); free(H); free(g); free(y); free(c); free(s); free(Ax); free(xbase); free(w);
  return iter;
}

/* ======================================================================
 * Grid extension (highreg_stokes_extend_to_grid/main.cpp:969-1830)
 * ====================================================================== */

#define GAT(N1, i, j, k) ((((size_t)(i)) * (N1) + (j)) * (N1) + (k))
#define GIN(N, i, j, k) ((((size_t)(i) - 1) * ((N) - 1) + (j) - 1) * ((N) - 1) + (k) - 1)

/* DST-I along one axis of an n^3 array with the dense sine matrix S. */
static void orc_dst_axis(int n, const double* S, double* f, int axis) {
  const size_t n2 = (size_t)n * n;
  const size_t stride = axis == 2 ? 1 : axis == 1 ? (size_t)n : n2;
#pragma omp parallel
  {
    double* x = (double*)malloc(sizeof(double) * n);
    double* y = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (long l = 0; l < (long)n2; l++) {
      /* line base: the two other indices of l */
      int a = (int)(l / n), b = (int)(l % n);
      size_t base = axis == 2 ? ((size_t)a * n + b) * n : axis == 1 ? (size_t)a * n2 + b : (size_t)a * n + b;
      for (int t = 0; t < n; t++) x[t] = f[base + t * stride];
      for (int k = 0; k < n; k++) {
        double acc = 0.0;
        for (int j = 0; j < n; j++) acc += x[j] * S[(size_t)j * n + k];
        y[k] = acc;
      }
      for (int t = 0; t < n; t++) f[base + t * stride] = y[t];
    }
    free(x);
    free(y);
  }
}

/* fftw_DST_Poisson3d_FD (main.cpp:1423-1491): forward DST-I (FFTW RODFT00,
 * Y_k = 2 sum x_j sin(pi (j+1)(k+1)/(n+1))) on all axes, division by the
 * eigenvalue factor, the same transform again, scale 1/(2(n+1))^3. */
void orc_dst_poisson3d(int n, double* f) {
  double* S = (double*)malloc(sizeof(double) * n * n);
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) S[(size_t)j * n + k] = 2.0 * sin(ORC_PI * (j + 1.0) * (k + 1.0) / (n + 1.0));
  for (int ax = 0; ax < 3; ax++) orc_dst_axis(n, S, f, ax);
  double pi_over_n = ORC_PI / (n + 1);
  for (int i = 0; i < n; i++) {
    double c1 = cos((i + 1) * pi_over_n);
    double s1 = sin((i + 1) * pi_over_n / 2.0);
    s1 *= -2.0 * s1;
    for (int j = 0; j < n; j++) {
      double c2 = cos((j + 1) * pi_over_n);
      double s2 = sin((j + 1) * pi_over_n / 2.0);
      s2 *= -2.0 * s2;
      for (int k = 0; k < n; k++) {
        double c3 = cos((k + 1) * pi_over_n);
        double s3 = sin((k + 1) * pi_over_n / 2.0);
        s3 *= -2.0 * s3;
        size_t ind = ((size_t)i * n + j) * n + k;
        double lap7 = s1 + s2 + s3;
        double sum8 = c1 * c2 * c3;
        double factor = 2.0 / 3.0 * (2.0 * lap7 + sum8 - 1.0);
        if (fabs(factor) > 1.e-14) f[ind] /= factor;
        else f[ind] = 0.0;
      }
    }
  }
  for (int ax = 0; ax < 3; ax++) orc_dst_axis(n, S, f, ax);
  double scale = 1. / pow(2 * (n + 1), 3);
  for (size_t q = 0; q < (size_t)n * n * n; q++) f[q] *= scale;
  free(S);
}

/* extend_into_cube (main.cpp:1494-1563). */
void orc_extend_into_cube(int N, double L, double h, const double* g, double* G) {
  const size_t N1 = (size_t)N + 1;
  for (int i = 0; i <= N; i++) {
    double x = i * h / L;
    for (int j = 0; j <= N; j++) {
      double y = j * h / L;
      for (int k = 0; k <= N; k++) {
        double z = k * h / L;
        double F, E1, E2, E3, C;
        F = x * g[GAT(N1, N, j, k)] + (1.0 - x) * g[GAT(N1, 0, j, k)];
        F += y * g[GAT(N1, i, N, k)] + (1.0 - y) * g[GAT(N1, i, 0, k)];
        F += z * g[GAT(N1, i, j, N)] + (1.0 - z) * g[GAT(N1, i, j, 0)];
        E1 = (1.0 - y) * (1.0 - z) * g[GAT(N1, i, 0, 0)] + (1.0 - y) * z * g[GAT(N1, i, 0, N)]
           + y * (1.0 - z) * g[GAT(N1, i, N, 0)] + y * z * g[GAT(N1, i, N, N)];
        E2 = (1.0 - x) * (1.0 - z) * g[GAT(N1, 0, j, 0)] + (1.0 - x) * z * g[GAT(N1, 0, j, N)]
           + x * (1.0 - z) * g[GAT(N1, N, j, 0)] + x * z * g[GAT(N1, N, j, N)];
        E3 = (1.0 - x) * (1.0 - y) * g[GAT(N1, 0, 0, k)] + (1.0 - x) * y * g[GAT(N1, 0, N, k)]
           + x * (1.0 - y) * g[GAT(N1, N, 0, k)] + x * y * g[GAT(N1, N, N, k)];
        C = (1.0 - x) * (1.0 - y) * (1.0 - z) * g[GAT(N1, 0, 0, 0)] + (1.0 - x) * (1.0 - y) * z * g[GAT(N1, 0, 0, N)]
          + (1.0 - x) * y * (1.0 - z) * g[GAT(N1, 0, N, 0)] + (1.0 - x) * y * z * g[GAT(N1, 0, N, N)]
          + x * (1.0 - y) * (1.0 - z) * g[GAT(N1, N, 0, 0)] + x * (1.0 - y) * z * g[GAT(N1, N, 0, N)]
          + x * y * (1.0 - z) * g[GAT(N1, N, N, 0)] + x * y * z * g[GAT(N1, N, N, N)];
        G[GAT(N1, i, j, k)] = F - E1 - E2 - E3 + C;
      }
    }
  }
}

static double orc_lap4(int N, const double* G, int i, int j, int k) {
  const size_t N1 = (size_t)N + 1;
  double lap7 = G[GAT(N1, i + 1, j, k)] + G[GAT(N1, i - 1, j, k)] + G[GAT(N1, i, j + 1, k)] + G[GAT(N1, i, j - 1, k)]
              + G[GAT(N1, i, j, k + 1)] + G[GAT(N1, i, j, k - 1)] - 6.0 * G[GAT(N1, i, j, k)];
  double temp1 = G[GAT(N1, i + 1, j + 1, k + 1)] + G[GAT(N1, i + 1, j - 1, k + 1)]
               + G[GAT(N1, i - 1, j + 1, k + 1)] + G[GAT(N1, i - 1, j - 1, k + 1)] - 4.0 * G[GAT(N1, i, j, k)];
  double temp2 = G[GAT(N1, i + 1, j + 1, k - 1)] + G[GAT(N1, i + 1, j - 1, k - 1)]
               + G[GAT(N1, i - 1, j + 1, k - 1)] + G[GAT(N1, i - 1, j - 1, k - 1)] - 4.0 * G[GAT(N1, i, j, k)];
  return 2.0 / 3.0 * (lap7 + (temp1 + temp2) / 8.0);
}

/* lapl_4ord_inCube (main.cpp:1566-1595). */
void orc_lapl_in_cube(int N, const double* G, double* rhs) {
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++) rhs[GIN(N, i, j, k)] = -orc_lap4(N, G, i, j, k);
}

/* lapl_4ord_nearSurfc (main.cpp:1598-1637): band == 2 marks BAND2 nodes. */
void orc_lapl_near_surface(int N, const signed char* band, const double* u, double* rhs) {
  const size_t N1 = (size_t)N + 1;
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++)
        if (band[GAT(N1, i, j, k)] == 2) rhs[GIN(N, i, j, k)] += orc_lap4(N, u, i, j, k);
}

/* dp_dx_4ord / dp_dy_4ord / dp_dz_4ord (main.cpp:1641-1830). */
void orc_dp_4ord(int N, double h, int axis, const signed char* band, const unsigned char* outside,
                 const double* p, double* rhs) {
  const size_t N1 = (size_t)N + 1;
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++) {
        size_t q = GAT(N1, i, j, k);
        int out_band2 = outside[q] && band[q] != 2;
        if (!out_band2) continue;
        int di = axis == 0, dj = axis == 1, dk = axis == 2;
        int c = axis == 0 ? i : axis == 1 ? j : k;
#define PV(o) p[GAT(N1, i + (o) * di, j + (o) * dj, k + (o) * dk)]
        double d;
        if (c == 0 || c == 1)
          d = -(25.0 / 12.0) * PV(0) + 4.0 * PV(1) - 3.0 * PV(2) + (4.0 / 3.0) * PV(3) - 0.25 * PV(4);
        else if (c == N || c == (N - 1))
          d = (25.0 / 12.0) * PV(0) - 4.0 * PV(-1) + 3.0 * PV(-2) - (4.0 / 3.0) * PV(-3) + 0.25 * PV(-4);
        else {
          d = -PV(2) + 8.0 * PV(1) - 8.0 * PV(-1) + PV(-2);
          d /= 12.0;
        }
#undef PV
        rhs[GIN(N, i, j, k)] += d * h;
      }
}

/* solve_Poisson (main.cpp:969-1044). */
void orc_solve_poisson(int N, double L, double h, const signed char* band, const unsigned char* outside,
                       const double* p, double* u, int variable) {
  const size_t N1 = (size_t)N + 1;
  double* w = (double*)malloc(sizeof(double) * N1 * N1 * N1);
  size_t nin = (size_t)(N - 1) * (N - 1) * (N - 1);
  double* rhs = (double*)calloc(nin, sizeof(double));
  orc_extend_into_cube(N, L, h, u, w);
  orc_lapl_in_cube(N, w, rhs);
  orc_lapl_near_surface(N, band, u, rhs);
  if (variable >= 1 && variable <= 3) orc_dp_4ord(N, h, variable - 1, band, outside, p, rhs);
  orc_dst_poisson3d(N - 1, rhs);
  for (size_t q = 0; q < N1 * N1 * N1; q++) u[q] = w[q];
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++) u[GAT(N1, i, j, k)] += rhs[GIN(N, i, j, k)];
  free(w);
  free(rhs);
}
