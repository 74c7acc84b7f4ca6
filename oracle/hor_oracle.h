/* CPU oracle for the regularized Stokes boundary-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (OpenMP) restatement of the
 * reference algorithms, used as the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it; the
 * product library never calls it.
 *
 * Layout conventions (same as the product C-ABI in include/hor.h):
 *   sources  : ns points; sx, sn, sdens are 3*ns interleaved xyz; sa is ns.
 *   targets  : nt points; tx, tx0, tn0, tf0/tg0 are 3*nt interleaved; tb is nt.
 *   outputs  : 3*nt interleaved, the reference's SL[3*i+k] layout.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   kernels, reg coeffs  : reference build (oracle/_ref), mpmath closed forms
 *                          at 50 digits, lambda=0 identity, r->0 limits.
 *   direct sums          : reference build on seeded inputs; brute-force long
 *                          double sum on tiny inputs.
 *   treecode             : reference build (P=6, N0=1000, theta=0.6, the
 *                          reference's compile-time values); for other P/N0,
 *                          convergence to the direct sum as P grows, and the
 *                          barycentric identity on polynomial densities.
 *   stencil/interp       : reference build; exactness on degree-5 polynomials.
 *   grid extension       : the discrete Poisson residual (the solution satisfies
 *                          the 4th-order stencil equations), transfinite
 *                          exactness on face values and trilinear data,
 *                          stencil exactness on polynomials, O(h^4)
 *                          convergence of the harmonic extension (the
 *                          reference's main.cpp needs FFTW: unbuildable here).
 *   pressure             : reference build (stokes_pressure_7ord/_away);
 *                          Lapl_DL_7ord = the s2 of Stokes_SL_7ord; the
 *                          translating spheroid's exact pressure.
 *   matvec, RHS          : reference build (solve_IE_1step, eval_SL).
 *   GMRES                : reference build full solve at h=1/16; dense
 *                          numpy solve on a small system.
 */
#ifndef HOR_ORACLE_H
#define HOR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar kernels (highreg_stokes_2surf/utilities.cpp:163-337) ---- */
void orc_In_024(double lam, double I[3]);
void orc_reg_coeff(double lam, double c[3]);
void orc_sl_kernel(double r, double d, const double c[3], double* s1, double* s2);
void orc_sl_kernel_on(double r, double d, double* s1, double* s2);
void orc_dl_kernel(double r, double d, const double c[3], double* s2, double* s3);
void orc_dl_kernel_on(double r, double d, double* s3);

/* ---- direct sums (stokeslets.cpp, stresslets.cpp) ---- */
void orc_sl_off(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sf, int nt, const double* tx,
                const double* tb, const double* tn0, const double* tf0, double* out);
void orc_sl_on(double DEL, int ns, const double* sx, const double* sn,
               const double* sa, const double* sf, int nt, const double* tx,
               double* out);
void orc_dl_off(double DEL, int ns, const double* sx, const double* sn,
                const double* sa, const double* sg, int nt, const double* tx,
                const double* tb, const double* tx0, const double* tn0,
                const double* tg0, double* out);
void orc_dl_on(double DEL, int ns, const double* sx, const double* sn,
               const double* sa, const double* sg, int nt, const double* tx,
               double* out);

/* ---- regularized pressure (highreg_stokes_extend_to_grid/stokes.cpp:99-203) ----
 * out: nt scalars.  tf0/tn0: force and normal at each target's nearest point. */
void orc_lapl_dl_kernel(double r, double d, const double c[3], double* s2);
void orc_pressure_off(double DEL, int ns, const double* sx, const double* sn,
                      const double* sa, const double* sf, int nt, const double* tx,
                      const double* tb, const double* tn0, const double* tf0, double* out);
void orc_pressure_away(int ns, const double* sx, const double* sa, const double* sf, int nt,
                       const double* tx, double* out);

/* ---- KITC treecode (KITC.h, KITC.cpp) ----
 * ctr: centre of the root box [ctr-1-1e-8, ctr+1+1e-8]^3 (KITC.cpp:30-69).
 * P: barycentric degree, N0: leaf size, theta: MAC parameter. */
typedef struct {
  int P, N0;
  double theta;
  double ctr[3];
} orc_tree_params;

void orc_sl_on_tree(const orc_tree_params* tp, double DEL, int ns,
                    const double* sx, const double* sn, const double* sa,
                    const double* sf, int nt, const double* tx, double* out);
void orc_sl_off_tree(const orc_tree_params* tp, double DEL, int ns,
                     const double* sx, const double* sn, const double* sa,
                     const double* sf, int nt, const double* tx,
                     const double* tb, const double* tn0, const double* tf0,
                     double* out);
void orc_dl_on_tree(const orc_tree_params* tp, double DEL, int ns,
                    const double* sx, const double* sn, const double* sa,
                    const double* sg, int nt, const double* tx, double* out);
void orc_dl_off_tree(const orc_tree_params* tp, double DEL, int ns,
                     const double* sx, const double* sn, const double* sa,
                     const double* sg, int nt, const double* tx,
                     const double* tb, const double* tx0, const double* tn0,
                     const double* tg0, double* out);

/* Tree structure introspection (for checking the product's host tree
 * builder): returns the node count; when cap >= count fills per node
 * members[2], interval[6] (x0,x1,y0,y1,z0,z1), centre[3], MAC, parent, and
 * perm[ns] (the particle order: perm[k] = original index at slot k). */
int orc_tree_structure(const orc_tree_params* tp, int ns, const double* sx,
                       int cap, int* members, double* interval, double* centre,
                       double* mac, int* parent, int* perm);

/* Per-node moments of the SL (6 comps) or DL (12 comps) kind, in tree
 * order, (P+1)^3 values per component; node 0 (root) is left zero. */
int orc_tree_moments(const orc_tree_params* tp, int dl, int ns,
                     const double* sx, const double* sn, const double* sa,
                     const double* sdens, int cap_nodes, double* moments);

/* ---- stencil + interpolation (highreg_stokes_2surf/main.cpp:682-778, 916-1046) ----
 * ctr: centre of the source surface; i123 per target = (i1, i2, i3).
 * stencil (optional, may be null): 36 source indices per target, row-major
 * [k1][k2], 0 where unset (the reference's sentinel).  g0: 3*nt. */
void orc_stencil(const double ctr[3], double h, int ns, const double* sx,
                 int nt, const int* nrst, const double* tx0, const int* i123,
                 int* stencil);
void orc_interp_g0(const double ctr[3], double h, int ns, const double* sx,
                   const double* sg, int nt, const int* nrst, const double* tx0,
                   const int* i123, double* g0);

/* ---- two-surface problem (main.cpp:140-677) ---- */
typedef struct {
  int n;                   /* quadrature points = targets on this surface */
  const double* x;         /* 3n */
  const double* nrml;      /* 3n */
  const double* area;      /* n  */
  double ctr[3];           /* surface centre */
  /* nearest data of this surface's points w.r.t. the OTHER surface: */
  const int* nrst;         /* n, 0 = far sentinel (surface.cpp:500-503) */
  const double* x0;        /* 3n */
  const double* n0;        /* 3n */
  const double* b;         /* n  */
  const int* i123;         /* 3n */
  const double* f;         /* 3n surface force (SL density), may be null */
  const double* f0;        /* 3n force at the nearest point on the other surface */
} orc_surface;

typedef struct {
  double h, DEL;
  int use_tree;
  orc_tree_params tree[2]; /* tree params of surface 1 and 2 (ctr = surface centre) */
  orc_surface s[2];
} orc_problem;

/* solve_IE_1step: Au = u - 2/3 [DL_11 + DL_21 ; DL_12 + DL_22](u). */
void orc_matvec(const orc_problem* pb, const double* u, double* Au);
/* eval_SL: sl = [SL_11 + SL_21 ; SL_12 + SL_22](f). */
void orc_eval_sl(const orc_problem* pb, double* sl);

/* GMRES(m) of main.cpp:210-416 on the operator of orc_matvec.
 * fix_restart=0 reproduces the reference (restart from v0=b, x rebuilt from
 * zero, main.cpp:286,411-414); 1 restarts from r=b-Ax.  Returns iterations;
 * resid_hist (may be null) receives iter_max values. */
int orc_gmres(const orc_problem* pb, int n, const double* b, double* x, int m,
              double tol, int iter_max, int fix_restart, double* resid_hist,
              double* final_resid);

/* ---- grid extension (highreg_stokes_extend_to_grid/main.cpp:969-1830) ----
 * Grid (N+1)^3, node (i,j,k) at ((i N1) + j) N1 + k, N1 = N+1; interior
 * unknowns ((i-1)(N-1) + j-1)(N-1) + k-1.  The DST-I is applied as the dense
 * sine matrix (O(n^4)), not through an FFT. */
void orc_dst_poisson3d(int n, double* f);
void orc_extend_into_cube(int N, double L, double h, const double* g, double* G);
void orc_lapl_in_cube(int N, const double* G, double* rhs);
void orc_lapl_near_surface(int N, const signed char* band, const double* u, double* rhs);
void orc_dp_4ord(int N, double h, int axis, const signed char* band,
                 const unsigned char* outside, const double* p, double* rhs);
void orc_solve_poisson(int N, double L, double h, const signed char* band,
                       const unsigned char* outside, const double* p, double* u,
                       int variable);

#ifdef __cplusplus
}
#endif
#endif
