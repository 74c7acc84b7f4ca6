// Example / check program for the C++ drop-in: reads surfaces and targets
// from a binary file, calls every function of hor_dropin.h and writes the
// results.
//
//   dropin_check <in.bin> <out.bin>
//
// in.bin (little endian): int32 N_quad, N_target, P, N0; float64 DEL, theta,
//   ctr[3]; per source x[3] f[3] g[3] Nrml[3] Area; per target x,y,z,
//   nrst_x,nrst_y,nrst_z, nrst_Nrml[3], nrst_f[3], nrst_g[3], b.
// out.bin: float64 blocks of 3*N_target each, in the order
//   SL_off, SL_on, DL_off, DL_on, SL_on_tree, SL_off_tree, DL_on_tree,
//   DL_off_tree, then 3 values per target for stokes_SL_7ord,
//   stokes_SL_7ord_onSurf, stokes_DL_7ord, stokes_DL_7ord_onSurf (first
//   min(N_target, 8) targets).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hor_dropin.h"

static void rd(FILE* f, void* p, size_t n) {
  if (std::fread(p, 1, n, f) != n) {
    std::fprintf(stderr, "short read\n");
    std::exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int hdr[4];
  double dh[5];
  rd(in, hdr, sizeof(hdr));
  rd(in, dh, sizeof(dh));
  int nq = hdr[0], nt = hdr[1];
  double DEL = dh[0];
  hor_kitc_set_params(hdr[2], hdr[3], dh[1]);
  hor_dropin_set_centre(1, dh + 2);
  std::vector<Surf_point> S(nq);
  for (auto& s : S) {
    double r[13];
    rd(in, r, sizeof(r));
    for (int k = 0; k < 3; k++) {
      s.x[k] = r[k];
      s.f[k] = r[3 + k];
      s.g[k] = r[6 + k];
      s.Nrml[k] = r[9 + k];
    }
    s.Area = r[12];
  }
  std::vector<Target_point> T(nt);
  for (auto& t : T) {
    double r[16];
    rd(in, r, sizeof(r));
    t.x = r[0]; t.y = r[1]; t.z = r[2];
    t.nrst_x = r[3]; t.nrst_y = r[4]; t.nrst_z = r[5];
    for (int k = 0; k < 3; k++) {
      t.nrst_Nrml[k] = r[6 + k];
      t.nrst_f[k] = r[9 + k];
      t.nrst_g[k] = r[12 + k];
    }
    t.b = r[15];
  }
  std::fclose(in);
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  std::vector<double> v(3 * (size_t)nt);
  auto put = [&](const std::vector<double>& a) { std::fwrite(a.data(), sizeof(double), a.size(), out); };
  eval_Stokes_SL_offSurf(0.0, DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_SL_onSurf(DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_DL_offSurf(0.0, DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_DL_onSurf(DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_SL_onSurf_tree(1, 0.0, DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_SL_offSurf_tree(1, 0.0, DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_DL_onSurf_tree(1, 0.0, DEL, nq, S, nt, T, v); put(v);
  eval_Stokes_DL_offSurf_tree(1, 0.0, DEL, nq, S, nt, T, v); put(v);
  std::vector<double> f(3 * (size_t)nq), g(3 * (size_t)nq);
  for (int j = 0; j < nq; j++)
    for (int k = 0; k < 3; k++) {
      f[3 * j + k] = S[j].f[k];
      g[3 * j + k] = S[j].g[k];
    }
  int m = nt < 8 ? nt : 8;
  std::vector<double> r(3);
  for (int i = 0; i < m; i++) {
    const Target_point& t = T[i];
    double f0n0 = t.nrst_f[0] * t.nrst_Nrml[0] + t.nrst_f[1] * t.nrst_Nrml[1] + t.nrst_f[2] * t.nrst_Nrml[2];
    std::vector<double> x0 = {t.nrst_x, t.nrst_y, t.nrst_z};
    stokes_SL_7ord(DEL, nq, S, f, f0n0, t.x, t.y, t.z, t.b, r); put(r);
    stokes_SL_7ord_onSurf(DEL, nq, S, f, f0n0, t.x, t.y, t.z, r); put(r);
    stokes_DL_7ord(DEL, nq, S, g, t.nrst_g, x0, t.nrst_Nrml, t.x, t.y, t.z, t.b, r); put(r);
    stokes_DL_7ord_onSurf(DEL, nq, S, g, t.nrst_g, t.x, t.y, t.z, r); put(r);
  }
  std::fclose(out);
  hor_dropin_release();
  return 0;
}
