// Test-only C-ABI harness around the unmodified reference sources of
// highreg_stokes_extend_to_grid (utilities.cpp, surface.cpp, stokes.cpp;
// main.cpp needs fftw3.h, which this image lacks, so it is not built).
// Built by oracle/Makefile into oracle/_ref/librefgrid.so.  Only tests/ and
// bench.py's CPU-baseline leg load it.

#include <cstring>
#include <vector>

#include "stokes.h"  // reference header: pulls surface.h, utilities.h

extern "C" {

// Geometry constants: ellipse a,b,c, ctr.
void refg_constants(double* out) {
  out[0] = ellipse_a; out[1] = ellipse_b; out[2] = ellipse_c; out[3] = ctr;
}

// Generate_Surface (surface.cpp:12-66).
int refg_generate_surface(int N, double h, int cap, double* x, double* nrml, double* area) {
  vector<Surf_point> s;
  Generate_Surface(N, h, &s);
  int n = (int)s.size();
  if (cap >= n) {
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < 3; k++) {
        x[3 * i + k] = s[i].x[k];
        nrml[3 * i + k] = s[i].Nrml[k];
      }
      area[i] = s[i].Area;
    }
  }
  return n;
}

// find_nrst_surfc (surface.cpp:303-392) on the grid [0,N]^3 with spacing h.
// Returns the number of band targets; fills (when cap >= count) the grid index
// ijk[3n], b[n], x0[3n], n0[3n], band2[n].
int refg_band_targets(int N, double h, double BAND1, double BAND2, int ns,
                      const double* sx, const double* sn, const double* sa, int cap,
                      int* ijk, double* b, double* x0, double* n0, int* band2) {
  vector<Surf_point> s(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) {
      s[i].x[k] = sx[3 * i + k];
      s[i].Nrml[k] = sn[3 * i + k];
    }
    s[i].Area = sa[i];
  }
  int*** nrst = new int**[N + 1];
  for (int i = 0; i <= N; i++) {
    nrst[i] = new int*[N + 1];
    for (int j = 0; j <= N; j++) nrst[i][j] = new int[N + 1];
  }
  vector<Target_point> T;
  find_nrst_surfc(N, h, BAND1, BAND2, ns, s, nrst, &T);
  int n = (int)T.size();
  if (cap >= n) {
    for (int i = 0; i <= N; i++)
      for (int j = 0; j <= N; j++)
        for (int k = 0; k <= N; k++) {
          int t = nrst[i][j][k];
          if (t >= 0) {
            ijk[3 * t] = i; ijk[3 * t + 1] = j; ijk[3 * t + 2] = k;
          }
        }
    for (int t = 0; t < n; t++) {
      b[t] = T[t].b;
      x0[3 * t] = T[t].nrst_x; x0[3 * t + 1] = T[t].nrst_y; x0[3 * t + 2] = T[t].nrst_z;
      for (int k = 0; k < 3; k++) n0[3 * t + k] = T[t].nrst_Nrml[k];
      band2[t] = T[t].band2 ? 1 : 0;
    }
  }
  for (int i = 0; i <= N; i++) {
    for (int j = 0; j <= N; j++) delete[] nrst[i][j];
    delete[] nrst[i];
  }
  delete[] nrst;
  return n;
}

// stokes_SL_7ord (stokes.cpp:53-97) at each target; f0n0[n] is f0.n0.
void refg_sl_7ord(double DEL, int ns, const double* sx, const double* sn,
                  const double* sa, const double* sf, int nt, const double* tx,
                  const double* tb, const double* f0n0, double* out) {
  vector<Surf_point> s(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) {
      s[i].x[k] = sx[3 * i + k];
      s[i].Nrml[k] = sn[3 * i + k];
    }
    s[i].Area = sa[i];
  }
  vector<double> f(sf, sf + 3 * ns);
#pragma omp parallel for
  for (int t = 0; t < nt; t++) {
    vector<double> SL(3, 0);
    stokes_SL_7ord(DEL, ns, s, f, f0n0[t], tx[3 * t], tx[3 * t + 1], tx[3 * t + 2], tb[t], SL);
    for (int k = 0; k < 3; k++) out[3 * t + k] = SL[k];
  }
}

// stokes_SL_away (stokes.cpp:13-45).
void refg_sl_away(int ns, const double* sx, const double* sa, const double* sf,
                  int nt, const double* tx, double* out) {
  vector<Surf_point> s(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) s[i].x[k] = sx[3 * i + k];
    s[i].Area = sa[i];
  }
  vector<double> f(sf, sf + 3 * ns);
#pragma omp parallel for
  for (int t = 0; t < nt; t++) {
    vector<double> SL(3, 0);
    stokes_SL_away(ns, s, f, tx[3 * t], tx[3 * t + 1], tx[3 * t + 2], SL);
    for (int k = 0; k < 3; k++) out[3 * t + k] = SL[k];
  }
}

// stokes_pressure_7ord (stokes.cpp:134-203) at each target: f0[3n] the force
// at the nearest point, n0[3n] its normal.
void refg_pressure_7ord(double DEL, int ns, const double* sx, const double* sn,
                        const double* sa, const double* sf, int nt, const double* tx,
                        const double* tb, const double* tn0, const double* tf0, double* out) {
  vector<Surf_point> s(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) {
      s[i].x[k] = sx[3 * i + k];
      s[i].Nrml[k] = sn[3 * i + k];
    }
    s[i].Area = sa[i];
  }
  vector<double> f(sf, sf + 3 * ns);
#pragma omp parallel for
  for (int t = 0; t < nt; t++) {
    vector<double> f0(tf0 + 3 * t, tf0 + 3 * t + 3), n0(tn0 + 3 * t, tn0 + 3 * t + 3);
    out[t] = stokes_pressure_7ord(DEL, ns, s, f, f0, n0, tx[3 * t], tx[3 * t + 1], tx[3 * t + 2], tb[t]);
  }
}

// stokes_pressure_away (stokes.cpp:102-127).
void refg_pressure_away(int ns, const double* sx, const double* sa, const double* sf,
                        int nt, const double* tx, double* out) {
  vector<Surf_point> s(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) s[i].x[k] = sx[3 * i + k];
    s[i].Area = sa[i];
  }
  vector<double> f(sf, sf + 3 * ns);
#pragma omp parallel for
  for (int t = 0; t < nt; t++) out[t] = stokes_pressure_away(ns, s, f, tx[3 * t], tx[3 * t + 1], tx[3 * t + 2]);
}

// p_exact (utilities.cpp:294-321).
void refg_p_exact(int n, const double* x, double* p) {
  for (int i = 0; i < n; i++) p[i] = p_exact(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
}

// stokeslet_density (utilities.cpp:267-285) and SL_exact (360-411).
void refg_density(int n, const double* x, double* f) {
  for (int i = 0; i < n; i++) {
    vector<double> p(x + 3 * i, x + 3 * i + 3);
    vector<double> r = stokeslet_density(p);
    for (int k = 0; k < 3; k++) f[3 * i + k] = r[k];
  }
}

void refg_sl_exact(int n, const double* x, double* u) {
  for (int i = 0; i < n; i++) {
    vector<double> r = SL_exact(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    for (int k = 0; k < 3; k++) u[3 * i + k] = r[k];
  }
}

}  // extern "C"
