// hor_dropin.h - the reference's C++ interface of the hot path, served by the
// GPU (libhor_dropin.so over libhor.so).
//
// Declares Surf_point / Target_point with the reference's members and layout
// (highreg_stokes_2surf/utilities.h:30-53) and the twelve sum functions with
// the reference's signatures:
//   stokeslets.h:9-29, stresslets.h:9-30, KITC.h:99-118.
// A program that includes the reference's own headers links against
// libhor_dropin.so unchanged (the mangled names are identical).  Output vectors
// must be pre-sized to 3*N_target, as with the reference.
//
// Behaviour beyond the reference:
//   - treecode parameters P, N0, theta are runtime (hor_kitc_set_params;
//     defaults 6, 1000, 0.6 = KITC.h:14-17) and the root-box centres of
//     surf_num 1 and 2 are settable (defaults = utilities.h:20-26);
//   - errors (no GPU, non-finite input) throw std::runtime_error instead of
//     printing; there is no CPU fallback;
//   - plans (sorted geometry, regularization coefficients, tree, interaction
//     lists) are cached by geometry content, so repeated calls on the same
//     surfaces (the GMRES matvec) only upload densities.
#ifndef HOR_DROPIN_H
#define HOR_DROPIN_H

#include <vector>

using std::vector;

struct Surf_point {
  Surf_point() : x(3, 0), f(3, 0), g(3, 0), Nrml(3, 0), Area(0), plane(0), left(0), right(0), up(0), down(0) {}
  vector<double> x;
  vector<double> f;
  vector<double> g;
  vector<double> Nrml;
  double Area;
  int plane, left, right, up, down;
};

struct Target_point {
  Target_point()
      : x(0), y(0), z(0), nrst_x(0), nrst_y(0), nrst_z(0), nrst_Nrml(3, 0), nrst_f(3, 0), nrst_g(3, 0),
        u(0), v(0), w(0), nrst(0), i1(0), i2(0), i3(0), b(0) {}
  double x, y, z;
  double nrst_x, nrst_y, nrst_z;
  vector<double> nrst_Nrml, nrst_f, nrst_g;
  double u, v, w;
  int nrst;
  int i1;
  int i2;
  int i3;
  double b;
};

// stokeslets.h:9-29
void eval_Stokes_SL_offSurf(double h, double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                            const vector<Target_point>& Target, vector<double>& SL);
void stokes_SL_7ord(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& f,
                    double f0_dot_n0, double x, double y, double z, double b, vector<double>& SL);
void eval_Stokes_SL_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                           const vector<Target_point>& Target, vector<double>& SL_vec);
void stokes_SL_7ord_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& f,
                           double f0_dot_n0, double x, double y, double z, vector<double>& SL);
// stresslets.h:9-30
void eval_Stokes_DL_offSurf(double h, double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                            vector<Target_point>& Target, vector<double>& DL_vec);
void stokes_DL_7ord(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& g,
                    const vector<double>& g0, const vector<double>& x0, const vector<double>& n0, double x,
                    double y, double z, double b, vector<double>& DL);
void eval_Stokes_DL_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, int N_target,
                           vector<Target_point>& Target, vector<double>& DL_vec);
void stokes_DL_7ord_onSurf(double DEL, int N_quad, const vector<Surf_point>& Surfc, const vector<double>& g,
                           const vector<double>& g0, double x, double y, double z, vector<double>& DL);
// KITC.h:99-118
void eval_Stokes_SL_onSurf_tree(int surf_num, double h, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                int N_target, const vector<Target_point>& Target, vector<double>& SL);
void eval_Stokes_SL_offSurf_tree(int surf_num, double h, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                 int N_target, const vector<Target_point>& Target, vector<double>& SL);
void eval_Stokes_DL_onSurf_tree(int surf_num, double h, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                int N_target, const vector<Target_point>& Target, vector<double>& DL);
void eval_Stokes_DL_offSurf_tree(int surf_num, double h, double DEL, int N_quad, const vector<Surf_point>& Surfc,
                                 int N_target, const vector<Target_point>& Target, vector<double>& SL);

// Side channel (no counterpart in the reference, where these are compile-time
// constants): treecode parameters and the root-box centre of surf_num 1 / 2.
extern "C" void hor_kitc_set_params(int P, int N0, double theta);
extern "C" void hor_dropin_set_centre(int surf_num, const double ctr[3]);
// Releases the cached plans and the GPU context.
extern "C" void hor_dropin_release(void);

#endif
