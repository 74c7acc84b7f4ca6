// Host geometry: the input producers of the hot path (run once per geometry).
//
// Quadrature by grid-line/surface intersection with partition-of-unity
// weights, nearest-point data of target points, surface forcing, band targets
// and the translating-body densities / exact velocity.  Each function follows
// the reference file:line given beside it (paths relative to the reference
// root) and keeps its arithmetic order where the result feeds a parity check.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "geom_core.h"
#include "hor_internal.h"

namespace hor {

static const double kPi = 3.14159265358979323846;

using namespace geom;

// surfc_force (surface.cpp:530-559), geom_core.h.
void surface_force(const hor_body& B, const double x0[3], const double n0[3], double f[3]) {
  geom::surface_force(B, x0, n0, f);
}

// Generate_Surface (surface.cpp:13-71) in the body-centred frame, then
// translated by the centre (highreg_stokes_2surf/main.cpp:89-98).
int surface_generate(const hor_body& B, double h, double pou_deg, std::vector<double>& X,
                     std::vector<double>& Nr, std::vector<double>& Ar, std::vector<int>& Pl) {
  double rmax = std::max(B.abc[0], std::max(B.abc[1], B.abc[2]));
  int N = (int)std::ceil(rmax / h) + 1;
  double theta = pou_deg * kPi / 180.0;
  double ct = std::cos(theta);
  double A[3] = {0, 0, 0}, Bp[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) {
    for (int l1 = -N; l1 <= N; l1++) {
      for (int l2 = -N; l2 <= N; l2++) {
        if (i == 0) { A[1] = l1 * h; A[2] = l2 * h; Bp[1] = l1 * h; Bp[2] = l2 * h; }
        if (i == 1) { A[0] = l1 * h; A[2] = l2 * h; Bp[0] = l1 * h; Bp[2] = l2 * h; }
        if (i == 2) { A[0] = l1 * h; A[1] = l2 * h; Bp[0] = l1 * h; Bp[1] = l2 * h; }
        for (int j = -N; j <= N; j++) {
          A[i] = j * h;
          Bp[i] = A[i] + h;
          if (sign_phi(B, A) * sign_phi(B, Bp) < 0) {
            double tau = newton(B, A, h, i, 0, 1);
            double x[3] = {A[0], A[1], A[2]};
            x[i] += tau * h;
            double n[3];
            for (int k = 0; k < 3; k++) n[k] = dphi(B, k, x, false);
            double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++) n[k] /= nn;
            if (std::fabs(n[i]) >= ct) {
              double area = h * h * part_unity(i, n, theta) / std::fabs(n[i]);
              for (int k = 0; k < 3; k++) {
                X.push_back(x[k] + B.ctr[k]);
                Nr.push_back(n[k]);
              }
              Ar.push_back(area);
              Pl.push_back(i);
            }
          }
        }
      }
    }
  }
  return (int)Ar.size();
}


// ---------------------------------------------------------------------------
// k-d tree nearest-point search with the reference's tie rule (the first
// index among equal distances wins, find_Nearest_grid surface.cpp:370-390).

struct PointTree {
  struct Node {
    double lo[3], hi[3];
    int begin, end;   // range in idx
    int left, right;  // children (-1 for a leaf)
  };
  std::vector<Node> nodes;
  std::vector<int> idx;
  const double* x = nullptr;

  void build(const double* pts, const std::vector<int>& ids) {
    x = pts;
    idx = ids;
    nodes.clear();
    nodes.reserve(2 * ids.size() / 8 + 2);
    if (!idx.empty()) make(0, (int)idx.size());
  }
  int make(int b, int e) {
    Node nd;
    for (int k = 0; k < 3; k++) { nd.lo[k] = 1e300; nd.hi[k] = -1e300; }
    for (int q = b; q < e; q++)
      for (int k = 0; k < 3; k++) {
        nd.lo[k] = std::min(nd.lo[k], x[3 * idx[q] + k]);
        nd.hi[k] = std::max(nd.hi[k], x[3 * idx[q] + k]);
      }
    nd.begin = b;
    nd.end = e;
    nd.left = nd.right = -1;
    int me = (int)nodes.size();
    nodes.push_back(nd);
    if (e - b > 16) {
      int ax = 0;
      for (int k = 1; k < 3; k++)
        if (nd.hi[k] - nd.lo[k] > nd.hi[ax] - nd.lo[ax]) ax = k;
      int mid = (b + e) / 2;
      std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                       [&](int a, int c) { return x[3 * a + ax] < x[3 * c + ax]; });
      int l = make(b, mid);
      int r = make(mid, e);
      nodes[me].left = l;
      nodes[me].right = r;
    }
    return me;
  }
  static double box_d2(const Node& nd, const double q[3]) {
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) {
      double g = std::max(std::max(nd.lo[k] - q[k], q[k] - nd.hi[k]), 0.0);
      d2 += g * g;
    }
    return d2;
  }
  void search(int ni, const double q[3], double& bd, int& best) const {
    const Node& nd = nodes[ni];
    if (box_d2(nd, q) > bd) return;
    if (nd.left < 0) {
      for (int p = nd.begin; p < nd.end; p++) {
        int i = idx[p];
        const double* s = x + 3 * i;
        double d = (q[0] - s[0]) * (q[0] - s[0]) + (q[1] - s[1]) * (q[1] - s[1]) +
                   (q[2] - s[2]) * (q[2] - s[2]);
        if (d < bd || (d == bd && best >= 0 && i < best)) { bd = d; best = i; }
      }
      return;
    }
    double dl = box_d2(nodes[nd.left], q), dr = box_d2(nodes[nd.right], q);
    if (dl <= dr) { search(nd.left, q, bd, best); search(nd.right, q, bd, best); }
    else { search(nd.right, q, bd, best); search(nd.left, q, bd, best); }
  }
  // Nearest point to q with squared distance < limit2; -1 if none.
  int nearest(const double q[3], double limit2, double* d2out) const {
    int best = -1;
    double bd = limit2;
    if (!nodes.empty()) search(0, q, bd, best);
    if (d2out) *d2out = bd;
    return best;
  }
};

// Compute_nearest_data (surface.cpp:450-526).
int nearest_data(const hor_body& B, double h, double DEL, int ns, const double* sx,
                 const int* splane, int nt, const double* tx, int* nrst, double* x0,
                 double* n0, double* b, int* i123, double* f0) {
  PointTree grids[3];
  for (int p = 0; p < 3; p++) {
    std::vector<int> ids;
    for (int i = 0; i < ns; i++)
      if (splane[i] == p) ids.push_back(i);
    grids[p].build(sx, ids);
  }
  int failures = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : failures)
  for (int i = 0; i < nt; i++) {
    const double* pt = tx + 3 * i;
    if (f0) f0[3 * i] = f0[3 * i + 1] = f0[3 * i + 2] = 0.0;
    int i1, i2, i3;
    monge_patch(B, pt, i1, i2, i3);
    int ni = grids[i3].nearest(pt, 100.0, nullptr);  // min_dist = 100 (surface.cpp:374)
    nrst[i] = ni;
    if (ni < 0) { failures++; continue; }
    double xx0[3], nn0[3];
    if (nearest_point(B, pt, sx + 3 * ni, xx0, nn0) == -1) { failures++; continue; }
    for (int k = 0; k < 3; k++) {
      x0[3 * i + k] = xx0[k];
      n0[3 * i + k] = nn0[k];
    }
    i123[3 * i] = i1;
    i123[3 * i + 1] = i2;
    i123[3 * i + 2] = i3;
    b[i] = (pt[0] - xx0[0]) * nn0[0] + (pt[1] - xx0[1]) * nn0[1] + (pt[2] - xx0[2]) * nn0[2];
    if (std::fabs(b[i]) / DEL > 8.0) {
      nrst[i] = 0;  // far sentinel (surface.cpp:500-503)
      continue;
    }
    if (f0) geom::surface_force(B, xx0, nn0, f0 + 3 * i);
  }
  return failures;
}

// Band targets: grid nodes with |b| <= band, the set find_nrst_surfc
// (extend_to_grid/surface.cpp:303-392) tags with BAND1*h = band.  Candidates
// are pre-filtered by the level set (|x|_abc - 1 bounds the distance by
// a_min); the Newton projection (as find_Nearest) starts from the quadrature
// point nearest to the node's radial projection onto the body.  Nodes come out
// in (i, j, k) lexicographic order.
int band_targets(const hor_body& B, const double* origin, double h, int ngrid, double band,
                 int ns, const double* sx, std::vector<int>& ijk, std::vector<double>& X,
                 std::vector<double>& b, std::vector<double>& X0, std::vector<double>& N0) {
  const double tol = 1.e-12;
  double amin = std::min(B.abc[0], std::min(B.abc[1], B.abc[2]));
  double hi_bound = (1.0 + (band + 2 * h) / amin) * (1.0 + (band + 2 * h) / amin) - 1.0;
  double lo_r = 1.0 - (band + 2 * h) / amin;
  double lo_bound = lo_r > 0 ? lo_r * lo_r - 1.0 : -2.0;
  std::vector<int> all(ns);
  for (int i = 0; i < ns; i++) all[i] = i;
  PointTree tree;
  tree.build(sx, all);
  int n1 = ngrid + 1;
  std::vector<std::vector<int>> per_i(n1);
  std::vector<std::vector<double>> per_d(n1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n1; i++) {
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n1; k++) {
        double pt[3] = {origin[0] + i * h, origin[1] + j * h, origin[2] + k * h};
        double ph = phi(B, pt, true);
        if (ph > hi_bound || ph < lo_bound) continue;
        // radial projection onto the body: x_s = ctr + (pt - ctr)/|x_body|_abc
        double s = std::sqrt(ph + 1.0);
        if (!(s > 0)) continue;
        double xs[3];
        for (int d = 0; d < 3; d++) xs[d] = B.ctr[d] + (pt[d] - B.ctr[d]) / s;
        int q = tree.nearest(xs, 1e300, nullptr);
        if (q < 0) continue;
        double x0[3], n0[3];
        if (nearest_point(B, pt, sx + 3 * q, x0, n0) != 0) continue;
        double bb = (pt[0] - x0[0]) * n0[0] + (pt[1] - x0[1]) * n0[1] + (pt[2] - x0[2]) * n0[2];
        if (std::fabs(bb) <= band + tol) {
          per_i[i].push_back(j);
          per_i[i].push_back(k);
          double rec[7] = {bb, x0[0], x0[1], x0[2], n0[0], n0[1], n0[2]};
          per_d[i].insert(per_d[i].end(), rec, rec + 7);
        }
      }
  }
  for (int i = 0; i < n1; i++) {
    size_t m = per_i[i].size() / 2;
    for (size_t q = 0; q < m; q++) {
      int j = per_i[i][2 * q], k = per_i[i][2 * q + 1];
      ijk.push_back(i); ijk.push_back(j); ijk.push_back(k);
      X.push_back(origin[0] + i * h);
      X.push_back(origin[1] + j * h);
      X.push_back(origin[2] + k * h);
      const double* r = &per_d[i][7 * q];
      b.push_back(r[0]);
      X0.push_back(r[1]); X0.push_back(r[2]); X0.push_back(r[3]);
      N0.push_back(r[4]); N0.push_back(r[5]); N0.push_back(r[6]);
    }
  }
  return (int)b.size();
}

// ---------------------------------------------------------------------------
// Translating ellipsoid.  The single-layer density that reproduces a rigid
// translation U inside the body is f = F / (4 pi a b c) / sqrt(x^2/a^4 +
// y^2/b^4 + z^2/c^4) (body frame) with the Oberbeck drag
// F_i = 16 pi U_i / (chi + a_i^2 alpha_i), chi = int_0^inf dl/Delta,
// alpha_i = int_0^inf dl/((a_i^2+l) Delta), Delta = sqrt(prod (a_k^2+l)).
// The integrals are Carlson's symmetric forms: chi = 2 R_F(a^2,b^2,c^2),
// alpha_1 = (2/3) R_D(b^2, c^2, a^2) (B. C. Carlson, Numer. Algorithms 10
// (1995) 13-26, duplication algorithm).  For a prolate spheroid this is the
// reference's stokeslet_density (extend_to_grid/utilities.cpp:267-285).

static double carlson_rf(double x, double y, double z) {
  for (int it = 0; it < 100; it++) {
    double sx = std::sqrt(x), sy = std::sqrt(y), sz = std::sqrt(z);
    double l = sx * sy + sx * sz + sy * sz;
    x = 0.25 * (x + l); y = 0.25 * (y + l); z = 0.25 * (z + l);
    double m = (x + y + z) / 3.0;
    double dx = 1 - x / m, dy = 1 - y / m, dz = 1 - z / m;
    double e = std::max(std::fabs(dx), std::max(std::fabs(dy), std::fabs(dz)));
    if (e < 1e-4) {
      double e2 = dx * dy - dz * dz, e3 = dx * dy * dz;
      return (1 - e2 / 10 + e3 / 14 + e2 * e2 / 24 - 3 * e2 * e3 / 44) / std::sqrt(m);
    }
  }
  return NAN;
}

static double carlson_rd(double x, double y, double z) {
  double sum = 0.0, fac = 1.0;
  for (int it = 0; it < 100; it++) {
    double sx = std::sqrt(x), sy = std::sqrt(y), sz = std::sqrt(z);
    double l = sx * sy + sx * sz + sy * sz;
    sum += fac / (sz * (z + l));
    fac *= 0.25;
    x = 0.25 * (x + l); y = 0.25 * (y + l); z = 0.25 * (z + l);
    double m = (x + y + 3 * z) / 5.0;
    double dx = 1 - x / m, dy = 1 - y / m, dz = 1 - z / m;
    double e = std::max(std::fabs(dx), std::max(std::fabs(dy), std::fabs(dz)));
    if (e < 1e-4) {
      double ea = dx * dy, eb = dz * dz, ec = ea - eb, ed = ea - 6 * eb, ee = ed + ec + ec;
      double s = 1 + ed * (-3.0 / 14 + 9.0 / 88 * ed - 4.5 / 26 * dz * ee) +
                 dz * (ee / 6 + dz * (-9.0 / 22 * ec + 3.0 / 26 * dz * ea));
      return 3 * sum + fac * s / (m * std::sqrt(m));
    }
  }
  return NAN;
}

void traction_translating(const hor_body& B, const double U[3], int n, const double* x, double* f) {
  double a2[3] = {B.abc[0] * B.abc[0], B.abc[1] * B.abc[1], B.abc[2] * B.abc[2]};
  double chi = 2.0 * carlson_rf(a2[0], a2[1], a2[2]);
  double alpha[3] = {2.0 / 3.0 * carlson_rd(a2[1], a2[2], a2[0]),
                     2.0 / 3.0 * carlson_rd(a2[2], a2[0], a2[1]),
                     2.0 / 3.0 * carlson_rd(a2[0], a2[1], a2[2])};
  // U in the body frame: U_b = R U (R maps global to body directions).
  double Ub[3];
  if (B.rotated) {
    for (int i = 0; i < 3; i++) Ub[i] = B.R[3 * i] * U[0] + B.R[3 * i + 1] * U[1] + B.R[3 * i + 2] * U[2];
  } else {
    Ub[0] = U[0]; Ub[1] = U[1]; Ub[2] = U[2];
  }
  double Fb[3];
  for (int i = 0; i < 3; i++) Fb[i] = 16.0 * kPi * Ub[i] / (chi + a2[i] * alpha[i]);
  double F[3];
  if (B.rotated) {
    for (int i = 0; i < 3; i++) F[i] = B.R[i] * Fb[0] + B.R[3 + i] * Fb[1] + B.R[6 + i] * Fb[2];
  } else {
    F[0] = Fb[0]; F[1] = Fb[1]; F[2] = Fb[2];
  }
  double s = 1.0 / (4.0 * kPi * B.abc[0] * B.abc[1] * B.abc[2]);
  for (int p = 0; p < n; p++) {
    double u[3];
    to_body(B, x + 3 * p, true, u);
    double q = u[0] * u[0] / (a2[0] * a2[0]) + u[1] * u[1] / (a2[1] * a2[1]) + u[2] * u[2] / (a2[2] * a2[2]);
    double w = s / std::sqrt(q);
    for (int k = 0; k < 3; k++) f[3 * p + k] = F[k] * w;
  }
}

// Exact Stokes flow of an ellipsoid with semi-axes a_k translating with
// velocity U through fluid at rest at infinity (Oberbeck's solution), in the
// body frame.  With lambda(x) >= 0 the ellipsoidal coordinate of the
// confocal ellipsoid through x (sum_k x_k^2/(a_k^2 + lambda) = 1),
// Delta(l) = sqrt(prod_k (a_k^2 + l)),
//   chi(l)     = int_l^inf ds/Delta(s)                 = 2 R_F(a^2+l, b^2+l, c^2+l),
//   alpha_j(l) = int_l^inf ds/((a_j^2+s) Delta(s))     = (2/3) R_D(.., .., a_j^2+l),
// the velocity is the Papkovich-Neuber field v = grad(x.Phi + phi0) - 2 Phi
// with the harmonic potentials Phi = sum_j C_j chi e_j and
// phi0 = sum_j D_j dOmega/dx_j, Omega the exterior potential of the solid
// ellipsoid (dOmega/dx_j = -2 x_j alpha_j), D_j = C_j a_j^2/2 (no tangential
// slip on the surface) and C_j = -U_j/(chi(0) + a_j^2 alpha_j(0)) (v = U on it):
//   v_j = -C_j (chi + a_j^2 alpha_j) - lambda lambda_j/Delta sum_i C_i x_i/(a_i^2 + lambda),
//   lambda_j = dlambda/dx_j = (2 x_j/(a_j^2+lambda)) / sum_k x_k^2/(a_k^2+lambda)^2.
// For a sphere it is Stokes' solution, for a prolate spheroid translating
// along its axis SL_exact (extend_to_grid/utilities.cpp:360-411); its drag
// is F_j = -16 pi mu C_j = 16 pi mu U_j/(chi(0) + a_j^2 alpha_j(0)), the
// force of traction_translating's density.  Inside and on the body v = U.
void ellipsoid_exact(const hor_body& B, const double U[3], int n, const double* xin, double* uout) {
  const double a2[3] = {B.abc[0] * B.abc[0], B.abc[1] * B.abc[1], B.abc[2] * B.abc[2]};
  double Ub[3];
  if (B.rotated) {
    for (int i = 0; i < 3; i++) Ub[i] = B.R[3 * i] * U[0] + B.R[3 * i + 1] * U[1] + B.R[3 * i + 2] * U[2];
  } else {
    Ub[0] = U[0]; Ub[1] = U[1]; Ub[2] = U[2];
  }
  const double chi0 = 2.0 * carlson_rf(a2[0], a2[1], a2[2]);
  double Cc[3];
  for (int j = 0; j < 3; j++) {
    const double al0 = 2.0 / 3.0 * carlson_rd(a2[(j + 1) % 3], a2[(j + 2) % 3], a2[j]);
    Cc[j] = -Ub[j] / (chi0 + a2[j] * al0);
  }
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; p++) {
    double x[3];
    to_body(B, xin + 3 * p, true, x);
    double v[3];
    const double q0 = x[0] * x[0] / a2[0] + x[1] * x[1] / a2[1] + x[2] * x[2] / a2[2];
    if (!(q0 > 1.0)) {
      for (int j = 0; j < 3; j++) v[j] = Ub[j];
    } else {
      // lambda: the root of F(l) = sum x_k^2/(a_k^2+l) - 1 (decreasing) in (0, |x|^2]:
      // Newton from the upper end, safeguarded by bisection
      double lo = 0.0, hi = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], l = hi;
      for (int it = 0; it < 200; it++) {
        double F = -1.0, dF = 0.0;
        for (int k = 0; k < 3; k++) {
          const double d = 1.0 / (a2[k] + l);
          F += x[k] * x[k] * d;
          dF -= x[k] * x[k] * d * d;
        }
        if (F > 0.0) lo = l;
        else hi = l;
        double ln = l - F / dF;
        if (!(ln > lo && ln < hi)) ln = 0.5 * (lo + hi);
        if (std::fabs(ln - l) <= 1e-16 * std::max(l, 1e-300) || hi - lo <= 1e-16 * hi) {
          l = ln;
          break;
        }
        l = ln;
      }
      double e[3], S2 = 0.0, Delta = 1.0;
      for (int k = 0; k < 3; k++) {
        e[k] = a2[k] + l;
        S2 += x[k] * x[k] / (e[k] * e[k]);
        Delta *= e[k];
      }
      Delta = std::sqrt(Delta);
      const double chi = 2.0 * carlson_rf(e[0], e[1], e[2]);
      double sum = 0.0;
      for (int i = 0; i < 3; i++) sum += Cc[i] * x[i] / e[i];
      for (int j = 0; j < 3; j++) {
        const double alj = 2.0 / 3.0 * carlson_rd(e[(j + 1) % 3], e[(j + 2) % 3], e[j]);
        const double lj = 2.0 * x[j] / e[j] / S2;
        v[j] = -Cc[j] * (chi + a2[j] * alj) - l * lj / Delta * sum;
      }
    }
    if (B.rotated) {
      for (int i = 0; i < 3; i++) uout[3 * p + i] = B.R[i] * v[0] + B.R[3 + i] * v[1] + B.R[6 + i] * v[2];
    } else {
      for (int i = 0; i < 3; i++) uout[3 * p + i] = v[i];
    }
  }
}

// SL_exact (extend_to_grid/utilities.cpp:360-411): exterior velocity of a
// prolate spheroid (semi-axes a > b = c along body x) translating with
// U = (1,0,0); inside and on the body the velocity is U.
void spheroid_exact(const hor_body& B, int n, const double* xin, double* uout) {
  double a = B.abc[0], bb = B.abc[1];
  double a2 = a * a, b2 = bb * bb;
  double c = std::sqrt(a2 - b2);
  double e = c / a;
  double Le = std::log((1.0 + e) / (1.0 - e));
  double a1 = e * e / ((1.0 + e * e) * Le - 2.0 * e);
  double b1 = a1 * (1.0 - e * e) / (2.0 * e * e);
  for (int p = 0; p < n; p++) {
    double u[3];
    to_body(B, xin + 3 * p, true, u);
    double x = u[0], y = u[1], z = u[2];
    double r2 = y * y + z * z;
    double R1 = std::sqrt((x + c) * (x + c) + r2);
    double R2 = std::sqrt((x - c) * (x - c) + r2);
    double R1c = R1 - (x + c), R2c = R2 - (x - c);
    double B10;
    if (r2 < 1.e-14) B10 = (x < 0) ? std::log((x - c) / (x + c)) : std::log((x + c) / (x - c));
    else B10 = std::log(R2c / R1c);
    double B30 = (x + c) / R1 - (x - c) / R2;
    double dB1dx = 1.0 / R1 - 1.0 / R2;
    double dB1dyz = 0.0;
    if (r2 > 1.e-14) dB1dyz = 1.0 / R2 / R2c - 1.0 / R1 / R1c;
    double dB3dyz = 1.0 / R2 - 1.0 / R1 + x * dB1dyz;
    double v[3];
    v[0] = 2.0 * a1 * B10 - a1 * B30 + 2.0 * b1 * ((x - c) / R2 - (x + c) / R1 + B10 + x * dB1dx);
    double Er = a1 * (1.0 / R2 - 1.0 / R1) + 2.0 * b1 * dB3dyz;
    v[1] = Er * y;
    v[2] = Er * z;
    // body frame -> global frame
    if (B.rotated) {
      for (int i = 0; i < 3; i++) uout[3 * p + i] = B.R[i] * v[0] + B.R[3 + i] * v[1] + B.R[6 + i] * v[2];
    } else {
      for (int i = 0; i < 3; i++) uout[3 * p + i] = v[i];
    }
  }
}

// p_exact (extend_to_grid/utilities.cpp:294-321): pressure of the prolate
// spheroid translating with U = (1,0,0) along its axis, p = -2 alpha1
// (1/R1 - 1/R2) (Chwang & Wu, eq. 29a), in the body frame.
void spheroid_exact_pressure(const hor_body& B, int n, const double* xin, double* pout) {
  double a_sq = B.abc[0] * B.abc[0], b_sq = B.abc[1] * B.abc[1];
  double c = std::sqrt(a_sq - b_sq);
  double e = c / B.abc[0];
  double Le = std::log((1.0 + e) / (1.0 - e));
  double a1 = e * e / ((1.0 + e * e) * Le - 2.0 * e);
  for (int p = 0; p < n; p++) {
    double u[3];
    to_body(B, xin + 3 * p, true, u);
    double x = u[0], y = u[1], z = u[2];
    double r_sq = y * y + z * z;
    double R1 = std::sqrt((x + c) * (x + c) + r_sq);
    double R2 = std::sqrt((x - c) * (x - c) + r_sq);
    double dB1dx = 1.0 / R1 - 1.0 / R2;
    pout[p] = -2.0 * a1 * dB1dx;
  }
}

}  // namespace hor

// ---------------------------------------------------------------------------
// C ABI wrappers

extern "C" {

int hor_body_init(hor_body* body, const double* abc, const double* ctr, const double* R) {
  if (!body || !abc) return hor::set_error(HOR_E_ARG, "hor_body_init: null argument");
  for (int k = 0; k < 3; k++) {
    body->abc[k] = abc[k];
    body->ctr[k] = ctr ? ctr[k] : 0.0;
    if (!(abc[k] > 0)) return hor::set_error(HOR_E_ARG, "hor_body_init: semi-axes must be > 0");
  }
  body->rotated = R ? 1 : 0;
  for (int k = 0; k < 9; k++) body->R[k] = R ? R[k] : ((k % 4 == 0) ? 1.0 : 0.0);
  return HOR_OK;
}

int hor_surface_generate(const hor_body* body, double h, double pou_deg, int cap, int* n,
                         double* x, double* nrml, double* area, int* plane) {
  if (!body || !n || !(h > 0)) return hor::set_error(HOR_E_ARG, "hor_surface_generate: bad argument");
  std::vector<double> X, Nr, Ar;
  std::vector<int> Pl;
  int cnt = hor::surface_generate(*body, h, pou_deg, X, Nr, Ar, Pl);
  *n = cnt;
  if (cap >= cnt && cnt > 0) {
    if (!x || !nrml || !area || !plane) return hor::set_error(HOR_E_ARG, "hor_surface_generate: null output");
    std::memcpy(x, X.data(), sizeof(double) * X.size());
    std::memcpy(nrml, Nr.data(), sizeof(double) * Nr.size());
    std::memcpy(area, Ar.data(), sizeof(double) * Ar.size());
    std::memcpy(plane, Pl.data(), sizeof(int) * Pl.size());
  }
  return HOR_OK;
}

int hor_nearest_data(const hor_body* body, double h, double DEL, int ns, const double* sx,
                     const int* splane, int nt, const double* tx, int* nrst, double* x0,
                     double* n0, double* b, int* i123, double* f0) {
  if (!body || !sx || !splane || !tx || !nrst || !x0 || !n0 || !b || !i123 || ns <= 0 || nt < 0)
    return hor::set_error(HOR_E_ARG, "hor_nearest_data: bad argument");
  int fails = hor::nearest_data(*body, h, DEL, ns, sx, splane, nt, tx, nrst, x0, n0, b, i123, f0);
  if (fails) return hor::set_error(HOR_E_ARG, "hor_nearest_data: nearest point not found for %d targets", fails);
  return HOR_OK;
}

int hor_surface_force(const hor_body* body, int n, const double* x, const double* nrml, double* f) {
  if (!body || !x || !nrml || !f || n < 0) return hor::set_error(HOR_E_ARG, "hor_surface_force: bad argument");
  for (int i = 0; i < n; i++) hor::surface_force(*body, x + 3 * i, nrml + 3 * i, f + 3 * i);
  return HOR_OK;
}

int hor_band_targets(const hor_body* body, const double* origin, double h, int ngrid, double band,
                     int ns, const double* sx, int cap, int* count, int* ijk, double* x,
                     double* b, double* x0, double* n0) {
  if (!body || !origin || !sx || !count || ns <= 0 || !(h > 0) || ngrid < 1)
    return hor::set_error(HOR_E_ARG, "hor_band_targets: bad argument");
  std::vector<int> IJK;
  std::vector<double> X, Bv, X0, N0v;
  int cnt = hor::band_targets(*body, origin, h, ngrid, band, ns, sx, IJK, X, Bv, X0, N0v);
  *count = cnt;
  if (cap >= cnt && cnt > 0) {
    if (!ijk || !x || !b || !x0 || !n0) return hor::set_error(HOR_E_ARG, "hor_band_targets: null output");
    std::memcpy(ijk, IJK.data(), sizeof(int) * IJK.size());
    std::memcpy(x, X.data(), sizeof(double) * X.size());
    std::memcpy(b, Bv.data(), sizeof(double) * Bv.size());
    std::memcpy(x0, X0.data(), sizeof(double) * X0.size());
    std::memcpy(n0, N0v.data(), sizeof(double) * N0v.size());
  }
  return HOR_OK;
}

int hor_traction_translating(const hor_body* body, const double* U, int n, const double* x, double* f) {
  if (!body || !U || !x || !f || n < 0) return hor::set_error(HOR_E_ARG, "hor_traction_translating: bad argument");
  hor::traction_translating(*body, U, n, x, f);
  return HOR_OK;
}

int hor_spheroid_exact_pressure(const hor_body* body, int n, const double* x, double* p) {
  if (!body || !x || !p || n < 0) return hor::set_error(HOR_E_ARG, "hor_spheroid_exact_pressure: bad argument");
  if (!(body->abc[0] > body->abc[1] && body->abc[1] == body->abc[2]))
    return hor::set_error(HOR_E_ARG, "hor_spheroid_exact_pressure: needs a prolate spheroid a > b = c");
  hor::spheroid_exact_pressure(*body, n, x, p);
  return HOR_OK;
}

// phi (utilities.cpp:58-86) at n points.
int hor_level_set(const hor_body* body, int n, const double* x, double* phi_out) {
  if (!body || !x || !phi_out || n < 0) return hor::set_error(HOR_E_ARG, "hor_level_set: bad argument");
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) phi_out[i] = hor::geom::phi(*body, x + 3 * i, true);
  return HOR_OK;
}

int hor_translating_exact_velocity(const hor_body* body, const double* U, int n, const double* x, double* u) {
  if (!body || !U || !x || !u || n < 0)
    return hor::set_error(HOR_E_ARG, "hor_translating_exact_velocity: bad argument");
  hor::ellipsoid_exact(*body, U, n, x, u);
  return HOR_OK;
}

int hor_spheroid_exact_velocity(const hor_body* body, int n, const double* x, double* u) {
  if (!body || !x || !u || n < 0) return hor::set_error(HOR_E_ARG, "hor_spheroid_exact_velocity: bad argument");
  if (!(body->abc[0] > body->abc[1] && body->abc[1] == body->abc[2]))
    return hor::set_error(HOR_E_ARG, "hor_spheroid_exact_velocity: needs a prolate spheroid a > b = c");
  hor::spheroid_exact(*body, n, x, u);
  return HOR_OK;
}

}  // extern "C"
