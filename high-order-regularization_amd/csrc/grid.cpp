// Grid extension of the grid program on the host (per the target design the
// Poisson step stays on the CPU; the boundary-integral values that feed it
// come from the GPU plans):
//
//   solve_Poisson          highreg_stokes_extend_to_grid/main.cpp:969-1044
//   fftw_DST_Poisson3d_FD  main.cpp:1423-1491  (DST-I here through our own FFT)
//   extend_into_cube       main.cpp:1494-1563  (transfinite interpolation)
//   lapl_4ord_inCube       main.cpp:1566-1595
//   lapl_4ord_nearSurfc    main.cpp:1598-1637
//   dp_d{x,y,z}_4ord       main.cpp:1641-1830
//
// Grid: nodes (i, j, k), 0 <= i, j, k <= N, at (i h, j h, k h), h = L / N,
// stored as u[(i (N+1) + j) (N+1) + k]; the Poisson unknowns are the (N-1)^3
// interior nodes, rhs[((i-1)(N-1) + j-1)(N-1) + k-1] as in the reference.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "hor_internal.h"

namespace hor {
namespace grid {

using cplx = std::complex<double>;
static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Complex FFT of length M = product of radices in {2, 3, 4, 5} (else none):
// decimation in time, recursive, each combine in place over the index set
// {k + m s}.  Forward sign: X_t = sum_m x_m exp(-2 pi i m t / M).

struct FFT {
  int M = 0;
  std::vector<int> radix;
  std::vector<cplx> W;  // W[t] = exp(-2 pi i t / M)

  bool plan(int m) {
    M = m;
    radix.clear();
    int r = m;
    while (r % 4 == 0) { radix.push_back(4); r /= 4; }
    while (r % 2 == 0) { radix.push_back(2); r /= 2; }
    while (r % 3 == 0) { radix.push_back(3); r /= 3; }
    while (r % 5 == 0) { radix.push_back(5); r /= 5; }
    if (r != 1) return false;
    W.resize(m);
    for (int t = 0; t < m; t++) W[t] = cplx(std::cos(2.0 * kPi * t / m), -std::sin(2.0 * kPi * t / m));
    return true;
  }

  // out[0..n) = DFT_n of in[0], in[stride], ...; level = index into radix.
  void rec(const cplx* in, cplx* out, int n, int stride, int level) const {
    if (n == 1) {
      out[0] = in[0];
      return;
    }
    const int p = radix[level], m = n / p;
    for (int q = 0; q < p; q++) rec(in + q * stride, out + q * m, m, stride * p, level + 1);
    const int step = M / n;  // W_n^e = W[e step]
    cplx t[5];
    for (int k = 0; k < m; k++) {
      for (int q = 0; q < p; q++) t[q] = out[q * m + k] * W[(size_t)q * k * step % M];
      if (p == 2) {
        out[k] = t[0] + t[1];
        out[k + m] = t[0] - t[1];
      } else if (p == 4) {
        // W_4 = -i
        const cplx a = t[0] + t[2], b = t[0] - t[2], c = t[1] + t[3];
        const cplx d = (t[1] - t[3]) * cplx(0.0, -1.0);
        out[k] = a + c;
        out[k + m] = b + d;
        out[k + 2 * m] = a - c;
        out[k + 3 * m] = b - d;
      } else {
        for (int s = 0; s < p; s++) {
          cplx acc = 0.0;
          for (int q = 0; q < p; q++) acc += t[q] * W[(size_t)q * s * m * step % M];
          out[k + s * m] = acc;
        }
      }
    }
  }
};

// DST-I of length n, unnormalised as FFTW's RODFT00:
//   Y_k = 2 sum_{j<n} x_j sin(pi (j+1)(k+1) / (n+1)).
// Two real lines per complex FFT of length 2(n+1) over the odd extension
// z = [0, x, 0, -reverse(x)]: the DFT of an odd real sequence is i times a
// real one, so Y1 = -Im(Z1), Y2 = -Im(Z2) come out of Z = Z1 + i Z2 as
// Y1_{k-1} = -Im Z_k and Y2_{k-1} = Re Z_k.  Lengths whose 2(n+1) has other
// prime factors use the sine matrix directly.
struct DST1 {
  int n = 0;
  bool fast = false;
  FFT fft;
  std::vector<double> S;  // n x n sine matrix when !fast

  void plan(int n_) {
    n = n_;
    fast = fft.plan(2 * (n + 1));
    if (!fast) {
      S.resize((size_t)n * n);
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) S[(size_t)j * n + k] = 2.0 * std::sin(kPi * (j + 1.0) * (k + 1.0) / (n + 1.0));
    }
  }

  // a, b: two lines of n values (b may be null), transformed in place.
  void run2(double* a, double* b, cplx* z, cplx* w) const {
    if (!fast) {
      std::vector<double> y(n);
      for (double* x : {a, b}) {
        if (!x) continue;
        for (int k = 0; k < n; k++) {
          double s = 0.0;
          for (int j = 0; j < n; j++) s += x[j] * S[(size_t)j * n + k];
          y[k] = s;
        }
        std::memcpy(x, y.data(), sizeof(double) * n);
      }
      return;
    }
    const int M = 2 * (n + 1);
    z[0] = 0.0;
    z[n + 1] = 0.0;
    for (int j = 0; j < n; j++) {
      const double xa = a[j], xb = b ? b[j] : 0.0;
      z[j + 1] = cplx(xa, xb);
      z[M - 1 - j] = cplx(-xa, -xb);
    }
    fft.rec(z, w, M, 1, 0);
    for (int k = 1; k <= n; k++) {
      a[k - 1] = -w[k].imag();
      if (b) b[k - 1] = w[k].real();
    }
  }
};

// In-place 3D DST-I of an n^3 row-major array along all three axes.
static void dst3(const DST1& D, double* f) {
  const int n = D.n;
  const size_t n2 = (size_t)n * n;
  const int M = 2 * (n + 1);
  // axis k (contiguous): line (i, j) at f + (i n + j) n
#pragma omp parallel
  {
    std::vector<cplx> z(M), w(M);
#pragma omp for schedule(static)
    for (long l = 0; l < (long)n2; l += 2) {
      double* a = f + (size_t)l * n;
      double* b = (l + 1 < (long)n2) ? f + (size_t)(l + 1) * n : nullptr;
      D.run2(a, b, z.data(), w.data());
    }
  }
  // axes j and i: gather two neighbouring lines (k, k+1) at a time
  for (int axis = 0; axis < 2; axis++) {
    const size_t stride = axis == 0 ? (size_t)n : n2;  // j-lines, then i-lines
#pragma omp parallel
    {
      std::vector<cplx> z(M), w(M);
      std::vector<double> la(n), lb(n);
#pragma omp for schedule(static) collapse(2)
      for (int o = 0; o < n; o++)
        for (int k = 0; k < n; k += 2) {
          // line base: axis j -> (i = o, k); axis i -> (j = o, k)
          const size_t base = axis == 0 ? (size_t)o * n2 + k : (size_t)o * n + k;
          const bool two = k + 1 < n;
          for (int t = 0; t < n; t++) {
            la[t] = f[base + t * stride];
            if (two) lb[t] = f[base + t * stride + 1];
          }
          D.run2(la.data(), two ? lb.data() : nullptr, z.data(), w.data());
          for (int t = 0; t < n; t++) {
            f[base + t * stride] = la[t];
            if (two) f[base + t * stride + 1] = lb[t];
          }
        }
    }
  }
}

// fftw_DST_Poisson3d_FD (main.cpp:1423-1491): f = h^2 rhs on the n^3 interior
// nodes in, the solution of the 4th-order (Iserles) discrete Poisson problem
// with zero Dirichlet data out.  Eigenvalue factor per mode
// (2/3)(2 (s1+s2+s3) + c1 c2 c3 - 1), s = -2 sin^2(theta/2), c = cos(theta),
// theta = pi (i+1)/(n+1); modes with |factor| <= 1e-14 are zeroed.
void dst_poisson3d(int n, double* f) {
  DST1 D;
  D.plan(n);
  dst3(D, f);
  const double pi_over_n = kPi / (n + 1);
  std::vector<double> c(n), s(n);
  for (int i = 0; i < n; i++) {
    c[i] = std::cos((i + 1) * pi_over_n);
    double si = std::sin((i + 1) * pi_over_n / 2.0);
    s[i] = -2.0 * si * si;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        const size_t ind = ((size_t)i * n + j) * n + k;
        const double lap7 = s[i] + s[j] + s[k];
        const double sum8 = c[i] * c[j] * c[k];
        const double factor = 2.0 / 3.0 * (2.0 * lap7 + sum8 - 1.0);
        f[ind] = std::fabs(factor) > 1.e-14 ? f[ind] / factor : 0.0;
      }
  dst3(D, f);
  const double scale = 1.0 / std::pow(2.0 * (n + 1), 3);
  const size_t n3 = (size_t)n * n * n;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)n3; q++) f[q] *= scale;
}

// ---------------------------------------------------------------------------

struct G3 {
  int N;
  size_t n1;
  explicit G3(int N_) : N(N_), n1((size_t)N_ + 1) {}
  size_t at(int i, int j, int k) const { return ((size_t)i * n1 + j) * n1 + k; }
  size_t in(int i, int j, int k) const { return (((size_t)i - 1) * (N - 1) + j - 1) * (N - 1) + k - 1; }
};

// extend_into_cube (main.cpp:1494-1563): transfinite interpolation of the face
// values of g into the cube, G = F - E1 - E2 - E3 + C.
void extend_into_cube(int N, double L, double h, const double* g, double* Gout) {
  const G3 A(N);
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= N; i++) {
    const double x = i * h / L;
    for (int j = 0; j <= N; j++) {
      const double y = j * h / L;
      for (int k = 0; k <= N; k++) {
        const double z = k * h / L;
        double F = x * g[A.at(N, j, k)] + (1.0 - x) * g[A.at(0, j, k)];
        F += y * g[A.at(i, N, k)] + (1.0 - y) * g[A.at(i, 0, k)];
        F += z * g[A.at(i, j, N)] + (1.0 - z) * g[A.at(i, j, 0)];
        const double E1 = (1.0 - y) * (1.0 - z) * g[A.at(i, 0, 0)] + (1.0 - y) * z * g[A.at(i, 0, N)] +
                          y * (1.0 - z) * g[A.at(i, N, 0)] + y * z * g[A.at(i, N, N)];
        const double E2 = (1.0 - x) * (1.0 - z) * g[A.at(0, j, 0)] + (1.0 - x) * z * g[A.at(0, j, N)] +
                          x * (1.0 - z) * g[A.at(N, j, 0)] + x * z * g[A.at(N, j, N)];
        const double E3 = (1.0 - x) * (1.0 - y) * g[A.at(0, 0, k)] + (1.0 - x) * y * g[A.at(0, N, k)] +
                          x * (1.0 - y) * g[A.at(N, 0, k)] + x * y * g[A.at(N, N, k)];
        const double C = (1.0 - x) * (1.0 - y) * (1.0 - z) * g[A.at(0, 0, 0)] +
                         (1.0 - x) * (1.0 - y) * z * g[A.at(0, 0, N)] +
                         (1.0 - x) * y * (1.0 - z) * g[A.at(0, N, 0)] + (1.0 - x) * y * z * g[A.at(0, N, N)] +
                         x * (1.0 - y) * (1.0 - z) * g[A.at(N, 0, 0)] + x * (1.0 - y) * z * g[A.at(N, 0, N)] +
                         x * y * (1.0 - z) * g[A.at(N, N, 0)] + x * y * z * g[A.at(N, N, N)];
        Gout[A.at(i, j, k)] = F - E1 - E2 - E3 + C;
      }
    }
  }
}

// (2/3)(lap7 + (corners of the cube (+-1,+-1,+-1) - 8 u)/8) at node (i,j,k):
// the 4th-order discrete Laplacian times h^2 (Iserles, p.169, ex. 8.11).
static inline double lap4(const G3& A, const double* u, int i, int j, int k) {
  const double c = u[A.at(i, j, k)];
  const double lap7 = u[A.at(i + 1, j, k)] + u[A.at(i - 1, j, k)] + u[A.at(i, j + 1, k)] + u[A.at(i, j - 1, k)] +
                      u[A.at(i, j, k + 1)] + u[A.at(i, j, k - 1)] - 6.0 * c;
  const double t1 = u[A.at(i + 1, j + 1, k + 1)] + u[A.at(i + 1, j - 1, k + 1)] + u[A.at(i - 1, j + 1, k + 1)] +
                    u[A.at(i - 1, j - 1, k + 1)] - 4.0 * c;
  const double t2 = u[A.at(i + 1, j + 1, k - 1)] + u[A.at(i + 1, j - 1, k - 1)] + u[A.at(i - 1, j + 1, k - 1)] +
                    u[A.at(i - 1, j - 1, k - 1)] - 4.0 * c;
  return 2.0 / 3.0 * (lap7 + (t1 + t2) / 8.0);
}

// lapl_4ord_inCube (main.cpp:1566-1595): rhs = -h^2 Delta_h G.
void lapl_in_cube(int N, const double* G, double* rhs) {
  const G3 A(N);
#pragma omp parallel for schedule(static)
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++) rhs[A.in(i, j, k)] = -lap4(A, G, i, j, k);
}

// lapl_4ord_nearSurfc (main.cpp:1598-1637): rhs += h^2 Delta_h u at band-2 nodes.
void lapl_near_surface(int N, const signed char* band, const double* u, double* rhs) {
  const G3 A(N);
#pragma omp parallel for schedule(static)
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++)
        if (band[A.at(i, j, k)] == 2) rhs[A.in(i, j, k)] += lap4(A, u, i, j, k);
}

// dp_d{x,y,z}_4ord (main.cpp:1641-1830): rhs += h dp/dx_axis (scaled by h^2)
// at nodes outside the body and outside band 2; 4th-order central
// differences, one-sided next to the faces.
void dp_4ord(int N, double h, int axis, const signed char* band, const unsigned char* outside, const double* p,
             double* rhs) {
  const G3 A(N);
#pragma omp parallel for schedule(static)
  for (int i = 1; i < N; i++)
    for (int j = 1; j < N; j++)
      for (int k = 1; k < N; k++) {
        const size_t q = A.at(i, j, k);
        if (!(outside[q] && band[q] != 2)) continue;
        const int idx[3] = {i, j, k};
        const int c = idx[axis];
        auto P = [&](int off) {
          int a[3] = {i, j, k};
          a[axis] += off;
          return p[A.at(a[0], a[1], a[2])];
        };
        double d;
        if (c == 0 || c == 1)
          d = -(25.0 / 12.0) * P(0) + 4.0 * P(1) - 3.0 * P(2) + (4.0 / 3.0) * P(3) - 0.25 * P(4);
        else if (c == N || c == N - 1)
          d = (25.0 / 12.0) * P(0) - 4.0 * P(-1) + 3.0 * P(-2) - (4.0 / 3.0) * P(-3) + 0.25 * P(-4);
        else
          d = (-P(2) + 8.0 * P(1) - 8.0 * P(-1) + P(-2)) / 12.0;
        rhs[A.in(i, j, k)] += d * h;
      }
}

// solve_Poisson (main.cpp:969-1044).
void solve_poisson(int N, double L, double h, const signed char* band, const unsigned char* outside, const double* p,
                   double* u, int variable) {
  const G3 A(N);
  std::vector<double> w(A.n1 * A.n1 * A.n1);
  extend_into_cube(N, L, h, u, w.data());
  std::vector<double> rhs((size_t)(N - 1) * (N - 1) * (N - 1), 0.0);
  lapl_in_cube(N, w.data(), rhs.data());
  lapl_near_surface(N, band, u, rhs.data());
  if (variable >= 1 && variable <= 3) dp_4ord(N, h, variable - 1, band, outside, p, rhs.data());
  dst_poisson3d(N - 1, rhs.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N; j++)
      for (int k = 0; k <= N; k++) {
        double v = w[A.at(i, j, k)];
        if (i > 0 && i < N && j > 0 && j < N && k > 0 && k < N) v += rhs[A.in(i, j, k)];
        u[A.at(i, j, k)] = v;
      }
}

}  // namespace grid
}  // namespace hor

// ---------------------------------------------------------------------------
// C ABI.

extern "C" int hor_dst_poisson3d(int n, double* f) {
  if (n < 1 || !f) return hor::set_error(HOR_E_ARG, "hor_dst_poisson3d: bad argument");
  hor::grid::dst_poisson3d(n, f);
  return HOR_OK;
}

extern "C" int hor_grid_solve_poisson(int N, double L, const signed char* band, const unsigned char* outside,
                                      const double* p, double* u, int variable) {
  if (N < 5 || !(L > 0) || !band || !outside || !u || variable < 0 || variable > 3 ||
      (variable >= 1 && !p))
    return hor::set_error(HOR_E_ARG, "hor_grid_solve_poisson: bad argument");
  hor::grid::solve_poisson(N, L, L / N, band, outside, p, u, variable);
  return HOR_OK;
}
