// Test-only C-ABI harness around the unmodified reference sources of
// highreg_stokes_2surf (compiled from /root/reference by oracle/Makefile into
// oracle/_ref/libref2surf.so).  It packs flat arrays into the reference's
// Surf_point / Target_point vectors, calls the reference function, and copies
// the result out.  Nothing here computes anything of its own.
//
// This file is part of the checker: only tests/ and bench.py's CPU-baseline leg
// load the library it builds.

#include <cstring>
#include <vector>

#include "stresslets.h"  // reference header: pulls stokeslets.h, KITC.h, surface.h, utilities.h

// Functions defined in the reference's main.cpp (no header there).
void interpolate_DL_density(int surf_num, double h, int N_quad,
                            const vector<Surf_point>& Surfc, int N_target,
                            vector<Target_point>& Target);
void solve_IE_1step(double h, double DEL_on, double DEL_off, int N_quad1,
                    vector<Surf_point>& Surfc1, int N_quad2,
                    vector<Surf_point>& Surfc2, int N_target1,
                    vector<Target_point>& Target1, int N_target2,
                    vector<Target_point>& Target2, int n,
                    const vector<double>& u_in, vector<double>& u_out);
void eval_SL(double h, double DEL_on, double DEL_off, int N_quad1,
             vector<Surf_point>& Surfc1, int N_quad2, vector<Surf_point>& Surfc2,
             int N_target1, vector<Target_point>& Target1, int N_target2,
             vector<Target_point>& Target2, vector<double>& SL_1,
             vector<double>& SL_2);
void test_Integral_Equation_gmres(double h, int N_quad1,
                                  vector<Surf_point>& Surfc1, int N_quad2,
                                  vector<Surf_point>& Surfc2, int N_target1,
                                  vector<Target_point>& Target1, int N_target2,
                                  vector<Target_point>& Target2);
void set_surface_force(int surf_num, int N_quad, vector<Surf_point>& Surfc);

namespace {

// Flat description of a surface: x, nrml, f, g are 3n interleaved; plane is n.
struct FlatSurf {
  int n;
  const double* x;
  const double* nrml;
  const double* area;
  const double* f;
  const double* g;
  const int* plane;
};

// Flat description of targets; any pointer may be null (then left at the
// reference's default-constructed value).
struct FlatTgt {
  int n;
  const double* x;      // 3n
  const double* b;      // n
  const double* x0;     // 3n  (nrst_x, nrst_y, nrst_z)
  const double* n0;     // 3n  (nrst_Nrml)
  const double* f0;     // 3n  (nrst_f)
  const double* g0;     // 3n  (nrst_g)
  const int* nrst;      // n
  const int* i123;      // 3n
};

vector<Surf_point> pack_surf(const FlatSurf& s) {
  vector<Surf_point> v(s.n);
  for (int i = 0; i < s.n; i++) {
    for (int k = 0; k < 3; k++) {
      v[i].x[k] = s.x[3 * i + k];
      if (s.nrml) v[i].Nrml[k] = s.nrml[3 * i + k];
      if (s.f) v[i].f[k] = s.f[3 * i + k];
      if (s.g) v[i].g[k] = s.g[3 * i + k];
    }
    if (s.area) v[i].Area = s.area[i];
    if (s.plane) v[i].plane = s.plane[i];
  }
  return v;
}

vector<Target_point> pack_tgt(const FlatTgt& t) {
  vector<Target_point> v(t.n);
  for (int i = 0; i < t.n; i++) {
    Target_point& p = v[i];
    p.x = t.x[3 * i];
    p.y = t.x[3 * i + 1];
    p.z = t.x[3 * i + 2];
    if (t.b) p.b = t.b[i];
    if (t.x0) {
      p.nrst_x = t.x0[3 * i];
      p.nrst_y = t.x0[3 * i + 1];
      p.nrst_z = t.x0[3 * i + 2];
    }
    for (int k = 0; k < 3; k++) {
      if (t.n0) p.nrst_Nrml[k] = t.n0[3 * i + k];
      if (t.f0) p.nrst_f[k] = t.f0[3 * i + k];
      if (t.g0) p.nrst_g[k] = t.g0[3 * i + k];
    }
    if (t.nrst) p.nrst = t.nrst[i];
    if (t.i123) {
      p.i1 = t.i123[3 * i];
      p.i2 = t.i123[3 * i + 1];
      p.i3 = t.i123[3 * i + 2];
    }
  }
  return v;
}

void copy_out(const vector<double>& v, double* out) {
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}

}  // namespace

extern "C" {

// Reference geometry constants: ellipse a,b,c then ctr1 xyz, ctr2 xyz, use_tree,
// KITC P, N0, sq_theta, PoU angle theta.
void ref2_constants(double* out) {
  out[0] = ellipse_a; out[1] = ellipse_b; out[2] = ellipse_c;
  out[3] = ctr1_x; out[4] = ctr1_y; out[5] = ctr1_z;
  out[6] = ctr2_x; out[7] = ctr2_y; out[8] = ctr2_z;
  out[9] = use_tree; out[10] = P; out[11] = N0; out[12] = sq_theta;
  out[13] = theta;
}

// Generate_Surface (surface.cpp:13-71); returns the point count and fills the
// arrays when cap >= count.  Points are NOT translated (main.cpp:89-98 does that).
int ref2_generate_surface(int rotated, int N, double h, int cap, double* x,
                          double* nrml, double* area, int* plane) {
  vector<Surf_point> s;
  Generate_Surface(0, rotated, N, h, &s);
  int n = (int)s.size();
  if (cap >= n) {
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < 3; k++) {
        x[3 * i + k] = s[i].x[k];
        nrml[3 * i + k] = s[i].Nrml[k];
      }
      area[i] = s[i].Area;
      plane[i] = s[i].plane;
    }
  }
  return n;
}

void ref2_reg_coeff(int n, const double* lam, double* c) {
  for (int i = 0; i < n; i++) get_reg_coeff_7ord(lam[i], c[3 * i], c[3 * i + 1], c[3 * i + 2]);
}

// kind: 0 SL(c), 1 SL onSurf, 2 DL(c), 3 DL onSurf.  out holds 2 values per r.
void ref2_kernel(int kind, int n, const double* r, double d, const double* c, double* out) {
  for (int i = 0; i < n; i++) {
    double a = 0, b = 0;
    if (kind == 0) Stokes_SL_7ord(r[i], d, c[0], c[1], c[2], a, b);
    if (kind == 1) Stokes_SL_7ord_onSurf(r[i], d, a, b);
    if (kind == 2) Stokes_DL_7ord(r[i], d, c[0], c[1], c[2], a, b);
    if (kind == 3) Stokes_DL_7ord_onSurf(r[i], d, b);
    out[2 * i] = a;
    out[2 * i + 1] = b;
  }
}

// Compute_nearest_data (surface.cpp:450-526) for targets tx against surface
// (surf_num, rotated).  Outputs: nrst[n], x0[3n], n0[3n], f0[3n], b[n], i123[3n].
void ref2_nearest(int surf_num, int rotated, double h, double DEL, int ns,
                  const double* sx, const double* sn, const double* sa,
                  const int* splane, int nt, const double* tx, int* nrst,
                  double* x0, double* n0, double* f0, double* b, int* i123) {
  FlatSurf S{ns, sx, sn, sa, nullptr, nullptr, splane};
  FlatTgt T{nt, tx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  vector<Surf_point> surf = pack_surf(S);
  vector<Target_point> tgt = pack_tgt(T);
  Compute_nearest_data(surf_num, rotated, h, DEL, ns, surf, nt, tgt);
  for (int i = 0; i < nt; i++) {
    nrst[i] = tgt[i].nrst;
    x0[3 * i] = tgt[i].nrst_x;
    x0[3 * i + 1] = tgt[i].nrst_y;
    x0[3 * i + 2] = tgt[i].nrst_z;
    for (int k = 0; k < 3; k++) {
      n0[3 * i + k] = tgt[i].nrst_Nrml[k];
      f0[3 * i + k] = tgt[i].nrst_f[k];
    }
    b[i] = tgt[i].b;
    i123[3 * i] = tgt[i].i1;
    i123[3 * i + 1] = tgt[i].i2;
    i123[3 * i + 2] = tgt[i].i3;
  }
}

// set_surface_force (main.cpp:1050-1059).
void ref2_surface_force(int surf_num, int n, const double* x, const double* nrml, double* f) {
  FlatSurf S{n, x, nrml, nullptr, nullptr, nullptr, nullptr};
  vector<Surf_point> surf = pack_surf(S);
  set_surface_force(surf_num, n, surf);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) f[3 * i + k] = surf[i].f[k];
}

// ---- direct sums (stokeslets.cpp, stresslets.cpp) -------------------------

void ref2_sl_off(double DEL, int ns, const double* sx, const double* sn,
                 const double* sa, const double* sf, int nt, const double* tx,
                 const double* tb, const double* tn0, const double* tf0, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, sf, nullptr, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, tb, nullptr, tn0, tf0, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_SL_offSurf(0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_sl_on(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sf, int nt, const double* tx, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, sf, nullptr, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_SL_onSurf(DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_dl_off(double DEL, int ns, const double* sx, const double* sn,
                 const double* sa, const double* sg, int nt, const double* tx,
                 const double* tb, const double* tx0, const double* tn0,
                 const double* tg0, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, nullptr, sg, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, tb, tx0, tn0, nullptr, tg0, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_DL_offSurf(0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_dl_on(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sg, int nt, const double* tx, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, nullptr, sg, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_DL_onSurf(DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

// ---- treecode (KITC.cpp:1389-1830); P, N0, theta are the reference's
// compile-time values (KITC.h:14-17) -----------------------------------------

void ref2_sl_on_tree(int surf_num, double DEL, int ns, const double* sx,
                     const double* sn, const double* sa, const double* sf, int nt,
                     const double* tx, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, sf, nullptr, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_SL_onSurf_tree(surf_num, 0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_sl_off_tree(int surf_num, double DEL, int ns, const double* sx,
                      const double* sn, const double* sa, const double* sf, int nt,
                      const double* tx, const double* tb, const double* tn0,
                      const double* tf0, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, sf, nullptr, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, tb, nullptr, tn0, tf0, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_SL_offSurf_tree(surf_num, 0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_dl_on_tree(int surf_num, double DEL, int ns, const double* sx,
                     const double* sn, const double* sa, const double* sg, int nt,
                     const double* tx, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, nullptr, sg, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_DL_onSurf_tree(surf_num, 0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

void ref2_dl_off_tree(int surf_num, double DEL, int ns, const double* sx,
                      const double* sn, const double* sa, const double* sg, int nt,
                      const double* tx, const double* tb, const double* tx0,
                      const double* tn0, const double* tg0, double* out) {
  vector<Surf_point> surf = pack_surf({ns, sx, sn, sa, nullptr, sg, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, tb, tx0, tn0, nullptr, tg0, nullptr, nullptr});
  vector<double> r(3 * nt);
  eval_Stokes_DL_offSurf_tree(surf_num, 0.0, DEL, ns, surf, nt, tgt, r);
  copy_out(r, out);
}

// interpolate_DL_density (main.cpp:916-1046): g0 at each target's nearest point.
void ref2_interp_g0(int surf_num, double h, int ns, const double* sx,
                    const double* sg, int nt, const double* tx, const int* tnrst,
                    const double* tx0, const double* tn0, const int* ti123,
                    double* g0_out) {
  vector<Surf_point> surf = pack_surf({ns, sx, nullptr, nullptr, nullptr, sg, nullptr});
  vector<Target_point> tgt = pack_tgt({nt, tx, nullptr, tx0, tn0, nullptr, nullptr, tnrst, ti123});
  interpolate_DL_density(surf_num, h, ns, surf, nt, tgt);
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) g0_out[3 * i + k] = tgt[i].nrst_g[k];
}

// Two-surface problem state for matvec / RHS / full solve.  Targets of surface
// k are the quadrature points of surface k (main.cpp:118-119); the nearest data
// of Target1 is relative to surface 2 and of Target2 relative to surface 1
// (main.cpp:165-166).
struct Ref2Problem {
  double h, DEL;
  vector<Surf_point> s1, s2;
  vector<Target_point> t1, t2;
};

static Ref2Problem g_prob;

// Sets up the problem.  nearest data arrays are for Target1 (w.r.t. surface 2)
// and Target2 (w.r.t. surface 1); pass null to have the reference compute them.
void ref2_problem_setup(double h, double DEL, int n1, const double* x1,
                        const double* nr1, const double* a1, const int* pl1,
                        const double* f1, int n2, const double* x2,
                        const double* nr2, const double* a2, const int* pl2,
                        const double* f2) {
  g_prob.h = h;
  g_prob.DEL = DEL;
  g_prob.s1 = pack_surf({n1, x1, nr1, a1, f1, nullptr, pl1});
  g_prob.s2 = pack_surf({n2, x2, nr2, a2, f2, nullptr, pl2});
  g_prob.t1 = pack_tgt({n1, x1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  g_prob.t2 = pack_tgt({n2, x2, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr});
  Compute_nearest_data(1, 0, h, DEL, n1, g_prob.s1, n2, g_prob.t2);
  Compute_nearest_data(2, 1, h, DEL, n2, g_prob.s2, n1, g_prob.t1);
}

// Copies out the nearest data the reference computed for target set k (1/2).
void ref2_problem_nearest(int k, int* nrst, double* x0, double* n0, double* f0,
                          double* b, int* i123) {
  const vector<Target_point>& t = (k == 1) ? g_prob.t1 : g_prob.t2;
  for (size_t i = 0; i < t.size(); i++) {
    nrst[i] = t[i].nrst;
    x0[3 * i] = t[i].nrst_x;
    x0[3 * i + 1] = t[i].nrst_y;
    x0[3 * i + 2] = t[i].nrst_z;
    for (int c = 0; c < 3; c++) {
      n0[3 * i + c] = t[i].nrst_Nrml[c];
      f0[3 * i + c] = t[i].nrst_f[c];
    }
    b[i] = t[i].b;
    i123[3 * i] = t[i].i1;
    i123[3 * i + 1] = t[i].i2;
    i123[3 * i + 2] = t[i].i3;
  }
}

// solve_IE_1step (main.cpp:541-677).
void ref2_problem_matvec(const double* u_in, double* u_out) {
  int n = 3 * (int)(g_prob.t1.size() + g_prob.t2.size());
  vector<double> in(u_in, u_in + n), out(n);
  solve_IE_1step(g_prob.h, g_prob.DEL, g_prob.DEL, (int)g_prob.s1.size(), g_prob.s1,
                 (int)g_prob.s2.size(), g_prob.s2, (int)g_prob.t1.size(), g_prob.t1,
                 (int)g_prob.t2.size(), g_prob.t2, n, in, out);
  copy_out(out, u_out);
}

// eval_SL (main.cpp:421-537) with the surface forces given at setup.
void ref2_problem_eval_sl(double* sl1, double* sl2) {
  vector<double> a(3 * g_prob.t1.size()), b(3 * g_prob.t2.size());
  eval_SL(g_prob.h, g_prob.DEL, g_prob.DEL, (int)g_prob.s1.size(), g_prob.s1,
          (int)g_prob.s2.size(), g_prob.s2, (int)g_prob.t1.size(), g_prob.t1,
          (int)g_prob.t2.size(), g_prob.t2, a, b);
  copy_out(a, sl1);
  copy_out(b, sl2);
}

// test_Integral_Equation_gmres (main.cpp:140-417): the whole solve.  It
// recomputes nearest data and forces itself and writes two velocity files to
// the current directory; the solution is read back from Target[i].u,v,w.
void ref2_full_solve(double h, double* u1, double* u2) {
  vector<Surf_point> s1 = g_prob.s1, s2 = g_prob.s2;
  vector<Target_point> a, b;
  Generate_Targets_OnSurface((int)s1.size(), s1, &a);
  Generate_Targets_OnSurface((int)s2.size(), s2, &b);
  test_Integral_Equation_gmres(h, (int)s1.size(), s1, (int)s2.size(), s2,
                               (int)a.size(), a, (int)b.size(), b);
  for (size_t i = 0; i < a.size(); i++) {
    u1[3 * i] = a[i].u; u1[3 * i + 1] = a[i].v; u1[3 * i + 2] = a[i].w;
  }
  for (size_t i = 0; i < b.size(); i++) {
    u2[3 * i] = b[i].u; u2[3 * i + 1] = b[i].v; u2[3 * i + 2] = b[i].w;
  }
}

}  // extern "C"
