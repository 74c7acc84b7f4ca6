// Error reporting, regularization coefficients and spatial ordering.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "hor_internal.h"

namespace hor {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// evaluate_In_024 (highreg_stokes_2surf/utilities.cpp:327-337) and
// get_reg_coeff_7ord (utilities.cpp:163-175).  lambda = b/delta; the
// moments I0, I2, I4 of the Gaussian tail give c3, c2, c1 in that order.
void reg_coeff_7ord(double lam, double c[3]) {
  const double rp = std::sqrt(3.14159265358979323846);
  double lam2 = lam * lam;
  double em2 = std::exp(-lam2);
  double t = std::fabs(lam) * std::erfc(std::fabs(lam));
  double I0 = em2 / rp - t;
  double I2 = 2.0 / 3.0 * ((0.5 - lam2) * em2 / rp + lam2 * t);
  double I4 = 8.0 / 15.0 * ((0.75 - 0.5 * lam2 + lam2 * lam2) * em2 / rp - lam2 * lam2 * t);
  double e2 = std::exp(lam2);
  double c3 = e2 * (2.0 * I0 - 4.0 * I2 + I4) / 8.0;
  double c2 = e2 * (I0 - I2) + (4.0 * lam2 + 7.0) * c3;
  double c1 = 2.0 * e2 * I0 + 2.0 * (lam2 + 1.0) * c2 - (4.0 * lam2 * lam2 + 6.0 * lam2 + 6.0) * c3;
  c[0] = c1;
  c[1] = c2;
  c[2] = c3;
}

static inline uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

uint64_t morton_key(const double p[3], const double lo[3], double ext) {
  uint64_t k = 0;
  for (int d = 0; d < 3; d++) {
    double s = (p[d] - lo[d]) / ext;
    s = std::min(std::max(s, 0.0), 1.0);
    uint64_t q = (uint64_t)(s * 2097151.0);
    k |= spread21(q) << d;
  }
  return k;
}

std::vector<int> spatial_order(int n, const double* x, const int* cls) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], x[3 * i + d]);
      hi[d] = std::max(hi[d], x[3 * i + d]);
    }
  double ext = 1e-300;
  for (int d = 0; d < 3; d++) ext = std::max(ext, hi[d] - lo[d]);
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; i++) key[i] = morton_key(x + 3 * i, lo, ext);
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    int ca = cls ? cls[a] : 0, cb = cls ? cls[b] : 0;
    if (ca != cb) return ca < cb;
    return key[a] < key[b];
  });
  return ord;
}

}  // namespace hor

extern "C" const char* hor_last_error(void) { return hor::g_err; }
extern "C" const char* hor_version(void) { return "hor-mi355x 0.1 (gfx950)"; }
