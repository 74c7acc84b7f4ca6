/* hor.h - C ABI of the MI355X regularized-Stokes library (libhor.so).
 *
 * The library evaluates the high-order regularized Stokeslet (single layer,
 * SL) and stresslet (double layer, DL) sums of Beale & Tlupova's method on one
 * AMD Instinct MI355X (gfx950), directly or with the barycentric-Lagrange
 * treecode (KITC), and drives the two-surface second-kind solve (matvec and
 * GMRES) with every vector resident in HBM.
 *
 * Conventions
 *   - All arrays are plain host pointers unless a name ends in _dev (a device
 *     pointer from hor_dev_alloc).  Points are float64 xyz interleaved (3*n),
 *     as the reference's Surf_point.x / Target_point and its SL[3*i+k] output.
 *   - Every function returns HOR_OK (0) or a negative HOR_E* code; the text of
 *     the last error on the calling thread is in hor_last_error().
 *   - Calls are synchronous: outputs are complete on return.
 *   - A context (hor_ctx) owns a HIP stream and device buffers; one context per
 *     host thread.  There is no global mutable state.
 *   - There is no CPU fallback: without a usable gfx950 device every compute
 *     entry point fails with HOR_E_NODEVICE.
 */
#ifndef HOR_H
#define HOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HOR_OK = 0,
  HOR_E_ARG = -1,       /* bad size / null pointer / bad enum */
  HOR_E_HIP = -2,       /* HIP runtime error */
  HOR_E_NODEVICE = -3,  /* no usable GPU */
  HOR_E_NAN = -4,       /* non-finite value detected */
  HOR_E_STATE = -5,     /* object used out of order */
  HOR_E_NOMEM = -6
};

/* Operator kinds. */
enum {
  HOR_SL_ON = 0,   /* eval_Stokes_SL_onSurf   (stokeslets.cpp:141-180) */
  HOR_SL_OFF = 1,  /* eval_Stokes_SL_offSurf  (stokeslets.cpp:12-86)   */
  HOR_DL_ON = 2,   /* eval_Stokes_DL_onSurf   (stresslets.cpp:186-228) */
  HOR_DL_OFF = 3,  /* eval_Stokes_DL_offSurf  (stresslets.cpp:13-96)   */
  HOR_P_OFF = 4,   /* stokes_pressure_7ord (extend_to_grid/stokes.cpp:134-203):
                      regularized pressure of the single layer with subtraction;
                      one scalar per target; direct method, FP64 only */
  HOR_SLDL_OFF = 5 /* eval_Stokes_SL_offSurf and eval_Stokes_DL_offSurf at the
                      same targets in one pass (the Stokeslet + stresslet
                      velocity of a grid band): the pairs share the
                      separation, 1/r and the erf/exp table; each kernel keeps
                      its own expressions, so the two results are those of
                      the SL_OFF and DL_OFF operators.  Direct method, FP64
                      or mixed; densities, nearest-point densities and results
                      are the SL and DL arrays back to back (see
                      hor_plan_eval) */
};

/* Modifier bits OR-ed into the kind of hor_plan_create (the reference's
 * per-target helpers called on their own):
 *   HOR_ALL_NEAR (SL_OFF / DL_OFF): every target takes the regularized branch
 *     whatever its |b|/delta, as stokes_SL_7ord / stokes_DL_7ord do
 *     (stokeslets.cpp:94-135, stresslets.cpp:104-179);
 *   HOR_ON_AT (DL_ON): the on-surface stresslet at arbitrary targets with the
 *     subtracted density g0 given per target in d0, as stokes_DL_7ord_onSurf
 *     (stresslets.cpp:236-279). */
enum { HOR_ALL_NEAR = 0x10, HOR_ON_AT = 0x20 };

/* Summation method. */
enum {
  HOR_DIRECT = 0,  /* O(N*M) direct sums (stokeslets.cpp / stresslets.cpp) */
  HOR_TREE = 1     /* KITC treecode (KITC.cpp) */
};

/* Arithmetic of the pair sums. */
enum {
  HOR_FP64 = 0,    /* every pair in FP64 (the parity path) */
  HOR_MIXED = 1    /* pairs with r/delta > 8 and treecode proxy interactions in
                      FP32, FP64 accumulation; near pairs and subtraction in FP64 */
};

typedef struct hor_ctx hor_ctx;

const char* hor_last_error(void);
const char* hor_version(void);

/* Number of visible devices (0 when none); does not create a context. */
int hor_device_count(int* count);

int hor_ctx_create(int device, hor_ctx** ctx);
int hor_ctx_destroy(hor_ctx* ctx);
/* Wall time of the last compute call on this context, in ms (HIP events on
 * the context's stream), and the kernel time of the dominant kernel. */
int hor_ctx_last_timing(hor_ctx* ctx, double* total_ms, double* kernel_ms);

/* Treecode parameters (KITC.h:14-17 are compile-time there; runtime here).
 * Defaults P=6, N0=1000, theta=0.6. */
typedef struct {
  int P;          /* barycentric degree, 1..12 */
  int N0;         /* leaf size */
  double theta;   /* MAC parameter */
} hor_tree_params;

/* ------------------------------------------------------------------------
 * Plans: geometry-only set-up is done once (AoS->SoA, spatial sort, per-target
 * regularization coefficients, treecode tree and interaction lists) and kept
 * on the device; densities change per call.
 *
 * A plan binds one source surface to one target set for one operator kind.
 *   sx, sn: 3*ns positions / unit normals; sa: ns quadrature weights.
 *   tx: 3*nt target positions (for *_ON, nt <= ns and target i is source i,
 *       as the reference assumes: stokeslets.cpp:167, KITC.cpp:1461,1678).
 *   tb: nt signed distances b (OFF only), tx0/tn0: 3*nt nearest surface point
 *       and its normal (OFF only; tx0 used by DL only).
 *   HOR_P_OFF takes the SL_OFF inputs; its result has one value per target.
 *   HOR_SLDL_OFF takes the DL_OFF inputs.
 *   ctr: root-box centre for HOR_TREE (KITC.cpp:45-60), ignored otherwise.
 * ------------------------------------------------------------------------ */
typedef struct hor_plan hor_plan;

int hor_plan_create(hor_ctx* ctx, int kind, int method, int precision,
                    double DEL, int ns, const double* sx, const double* sn,
                    const double* sa, int nt, const double* tx, const double* tb,
                    const double* tx0, const double* tn0, const double* ctr,
                    const hor_tree_params* tree, hor_plan** plan);
int hor_plan_destroy(hor_plan* plan);

/* Evaluates the plan's operator.
 *   dens: 3*ns source density (f for SL and P, g for DL), host.
 *   d0:   3*nt density at each target's nearest point (f0 = nrst_f for SL_OFF
 *         and P_OFF, g0 = nrst_g for DL_OFF), host; ignored for *_ON.
 *   out:  3*nt result (nt for P_OFF), host.
 * HOR_SLDL_OFF: dens = [f | g] (6*ns), d0 = [f0 | g0] (6*nt),
 *   out = [SL_OFF result | DL_OFF result] (6*nt). */
int hor_plan_eval(hor_plan* plan, const double* dens, const double* d0, double* out);

/* Device-resident variant: all three pointers are device pointers
 * (hor_dev_alloc); dens_dev 3*ns, d0_dev 3*nt or NULL, out_dev 3*nt (nt for
 * P_OFF; 6*ns, 6*nt and 6*nt for SLDL_OFF). */
int hor_plan_eval_dev(hor_plan* plan, const double* dens_dev, const double* d0_dev,
                      double* out_dev);

/* Pair interactions one evaluation performs (direct: nt*ns; tree: near pairs,
 * proxy interactions and moment terms separately). */
int hor_plan_stats(hor_plan* plan, double* direct_pairs, double* near_pairs,
                   double* proxy_pairs, double* moment_terms);

/* Far-field kernel of a mixed-precision single-layer treecode plan (the proxy
 * interactions of Call_Treecode_SL, KITC.cpp:671-720):
 *   HOR_FAR_VALU (default): FP32 vector arithmetic per proxy interaction;
 *   HOR_FAR_MFMA: the interactions as a GEMM on FP32 matrix cores,
 *     [targets x proxies] 1/r^3 times [proxies x 38] moment polynomials in
 *     node-centred coordinates, FP64 combination per target.
 * Any other plan accepts only HOR_FAR_VALU (HOR_E_ARG otherwise). */
enum { HOR_FAR_VALU = 0, HOR_FAR_MFMA = 1 };
int hor_plan_set_far_kernel(hor_plan* plan, int kernel);

/* Kernel time of each phase of a treecode plan's last evaluation, in ms (HIP
 * events on the context's stream): moments, near sums, far field (the
 * matrix-core far field includes building its B operand).  HOR_E_STATE
 * before the first evaluation; HOR_E_ARG for a direct plan. */
int hor_plan_phase_timing(hor_plan* plan, double* moments_ms, double* near_ms, double* far_ms);

/* ------------------------------------------------------------------------
 * One-shot drop-ins of the reference functions (plan, eval, destroy).
 * ------------------------------------------------------------------------ */
/* eval_Stokes_SL_offSurf (stokeslets.h:9-12); tn0/tf0 are nrst_Nrml / nrst_f. */
int hor_sl_offsurf(hor_ctx* ctx, double DEL, int ns, const double* sx,
                   const double* sn, const double* sa, const double* sf, int nt,
                   const double* tx, const double* tb, const double* tn0,
                   const double* tf0, double* out);
/* eval_Stokes_SL_onSurf (stokeslets.h:20-23). */
int hor_sl_onsurf(hor_ctx* ctx, double DEL, int ns, const double* sx,
                  const double* sn, const double* sa, const double* sf, int nt,
                  const double* tx, double* out);
/* eval_Stokes_DL_offSurf (stresslets.h:9-12); tx0/tn0/tg0 are nrst_x/y/z,
 * nrst_Nrml, nrst_g. */
int hor_dl_offsurf(hor_ctx* ctx, double DEL, int ns, const double* sx,
                   const double* sn, const double* sa, const double* sg, int nt,
                   const double* tx, const double* tb, const double* tx0,
                   const double* tn0, const double* tg0, double* out);
/* eval_Stokes_DL_onSurf (stresslets.h:21-24). */
int hor_dl_onsurf(hor_ctx* ctx, double DEL, int ns, const double* sx,
                  const double* sn, const double* sa, const double* sg, int nt,
                  const double* tx, double* out);
/* eval_Stokes_{SL,DL}_{on,off}Surf_tree (KITC.h:99-118); ctr = centre of the
 * source surface's root box (ctr1/ctr2 in the reference); tree may be NULL. */
int hor_sl_onsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree,
                       double DEL, int ns, const double* sx, const double* sn,
                       const double* sa, const double* sf, int nt,
                       const double* tx, double* out);
int hor_sl_offsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree,
                        double DEL, int ns, const double* sx, const double* sn,
                        const double* sa, const double* sf, int nt,
                        const double* tx, const double* tb, const double* tn0,
                        const double* tf0, double* out);
int hor_dl_onsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree,
                       double DEL, int ns, const double* sx, const double* sn,
                       const double* sa, const double* sg, int nt,
                       const double* tx, double* out);
int hor_dl_offsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree,
                        double DEL, int ns, const double* sx, const double* sn,
                        const double* sa, const double* sg, int nt,
                        const double* tx, const double* tb, const double* tx0,
                        const double* tn0, const double* tg0, double* out);

/* eval_Stokes_DL_onSurf (stresslets.h:21-24) as hor_dl_onsurf computes it,
 * with every pointer a device pointer (sx/sn/sg 3*ns, sa ns, tx 3*nt, out 3*nt;
 * target i is source i) and NO device memory allocated: no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * intermediate buffers live in one 32 MiB array of device memory that each GPU
 * holds from the library's first use until the process ends; sizes whose
 * buffers do not fit (above about ns = nt = 62000) are rejected with HOR_E_ARG.
 * Calls on one GPU share that array: calls that overlap on one GPU (from two
 * contexts or threads) race and must be serialized by the caller.  Sources and
 * targets keep the caller's order (no spatial sort), so the result equals
 * hor_dl_onsurf's up to the order of the sums.  Synchronous, like the rest of
 * the library. */
int hor_dl_onsurf_dev_noalloc(hor_ctx* ctx, double DEL, int ns, const double* sx_dev,
                              const double* sn_dev, const double* sa_dev,
                              const double* sg_dev, int nt, const double* tx_dev,
                              double* out_dev);

/* stokes_pressure_7ord (extend_to_grid/stokes.h:24-28) at nt targets: sf the
 * surface force, tb/tn0/tf0 the signed distance, nearest normal and the
 * force there; out: nt pressures. */
int hor_pressure_7ord(hor_ctx* ctx, double DEL, int ns, const double* sx,
                      const double* sn, const double* sa, const double* sf, int nt,
                      const double* tx, const double* tb, const double* tn0,
                      const double* tf0, double* out);
/* stokes_pressure_away (extend_to_grid/stokes.h:20-22): the plain sum
 * 1/(4 pi) sum A f.dx / r^3 at nt targets away from the surface. */
int hor_pressure_away(hor_ctx* ctx, int ns, const double* sx, const double* sn,
                      const double* sa, const double* sf, int nt, const double* tx,
                      double* out);

/* ------------------------------------------------------------------------
 * Two-surface second-kind solve (highreg_stokes_2surf/main.cpp:140-677).
 * Unknown order: [surface 1 xyz interleaved, surface 2 xyz interleaved].
 * For each surface k: quadrature x/nrml/area, centre ctr, and the nearest
 * data of its points relative to the OTHER surface (nrst: 0 = far sentinel,
 * x0, n0, b, i123 Monge indices), as Compute_nearest_data produces.
 * ------------------------------------------------------------------------ */
typedef struct {
  int n;
  const double* x;
  const double* nrml;
  const double* area;
  double ctr[3];
  const int* nrst;
  const double* x0;
  const double* n0;
  const double* b;
  const int* i123;
} hor_surface_desc;

typedef struct hor_solver hor_solver;

int hor_solver_create(hor_ctx* ctx, double h, double DEL, int method, int precision,
                      const hor_tree_params* tree1, const hor_tree_params* tree2,
                      const hor_surface_desc* s1, const hor_surface_desc* s2,
                      hor_solver** solver);
int hor_solver_destroy(hor_solver* solver);
/* n = 3*(n1+n2) */
int hor_solver_size(hor_solver* solver, int* n);
/* Work of one matvec summed over its four DL operators (see hor_plan_stats):
 * direct-sum-equivalent pairs, near (leaf or direct) pairs, proxy
 * interactions, moment terms. */
int hor_solver_stats(hor_solver* solver, double* direct_pairs, double* near_pairs,
                     double* proxy_pairs, double* moment_terms);

/* hor_solver_stats of one of the four DL plans: which = 0, 1 the on-surface
 * plans of surfaces 1 and 2 (DL11, DL22), 2, 3 the sources of surface 1 and 2
 * at the other surface's points (DL12, DL21). */
int hor_solver_plan_stats(hor_solver* solver, int which, double* direct_pairs,
                          double* near_pairs, double* proxy_pairs, double* moment_terms);

/* Kernel time (HIP events on the context's stream) of the phases of the last
 * matvec of a treecode solver's own operator, summed over its four DL plans:
 * the moments (k_moment_rhs, k_moments, k_moments_reduce), the near sums
 * (k_leaf) and the far field (k_far), in ms.  HOR_E_STATE for direct solvers. */
int hor_solver_phase_timing(hor_solver* solver, double* moments_ms, double* near_ms,
                            double* far_ms);

/* solve_IE_1step (main.cpp:541-677): Au = u - 2/3 [DL11+DL21; DL12+DL22](u),
 * with g0 interpolated on the 6x6 Monge stencil (main.cpp:916-1046).  Host
 * pointers; the _dev variant takes device pointers of length n. */
int hor_solver_matvec(hor_solver* solver, const double* u, double* Au);
int hor_solver_matvec_dev(hor_solver* solver, const double* u_dev, double* Au_dev);

/* eval_SL (main.cpp:421-537): sl = [SL11+SL21; SL12+SL22](f); f1/f2 are the
 * surface forces (3*n_k) and f01/f02 the force at each point's nearest point
 * on the other surface (nrst_f).  Host pointers. */
int hor_solver_eval_sl(hor_solver* solver, const double* f1, const double* f2,
                       const double* f01, const double* f02, double* sl);

/* Interpolated g0 (main.cpp:916-1046) at the points of surface k (1 or 2)
 * from the density u of the other surface; host pointers, out 3*n_k. */
int hor_solver_interp_g0(hor_solver* solver, int k, const double* u, double* g0);

/* Restarted GMRES(m) (main.cpp:210-416), device-resident, x0 = 0, restart from
 * the true residual b - Ax.  b, x host pointers of length n.  hist (may be
 * NULL) receives up to maxit relative residual estimates.  FP64 solver: stops
 * on the estimate and *relres is the final estimate |g_{j+1}|/|b|, as in the
 * reference.  Mixed-precision solver: the Arnoldi steps use the mixed operator,
 * every restart residual and the final check use the FP64 operator, a cycle
 * whose FP64 residual is above tol continues, and *relres is that FP64
 * residual, also when the solve stops at maxit.  A non-finite value in b, in a
 * residual or in an Arnoldi step ends the solve with HOR_E_NAN. */
int hor_solver_gmres(hor_solver* solver, const double* b, double* x, int m,
                     double tol, int maxit, int* iters, double* relres,
                     double* hist);

/* ------------------------------------------------------------------------
 * Learned right preconditioner.  No counterpart in the reference (its GMRES,
 * main.cpp:313-415, is unpreconditioned): a point-cloud message-passing
 * network (width 64, k nearest neighbours, 3 layers, BF16 matrix-core GEMMs
 * with FP32 accumulation) z = M(r) over the points of both surfaces, trained
 * on the GPU from (z, A z) pairs logged by hor_solver_fgmres.
 * ------------------------------------------------------------------------ */
typedef struct hor_mpnn hor_mpnn;

/* Number of FP32 parameters (fixed architecture). */
int hor_mpnn_param_count(void);
/* Network with k neighbours per point; parameters start at zero (M = I). */
int hor_mpnn_create(hor_ctx* ctx, int k, hor_mpnn** net);
int hor_mpnn_destroy(hor_mpnn* net);
/* Host FP32 parameter vectors of hor_mpnn_param_count() values; setting the
 * parameters resets the Adam state. */
int hor_mpnn_set_params(hor_mpnn* net, const float* params);
int hor_mpnn_get_params(hor_mpnn* net, float* params);
int hor_mpnn_get_grad(hor_mpnn* net, float* grad);
/* Point set in the solver's unknown order (n1 points of surface 1, then n2 of
 * surface 2; x, nrml 3*(n1+n2)), body centres ctr[6] and sizes size[2], and
 * the largest batch of vectors evaluated at once.  Builds the k-nearest-
 * neighbour graph on the host; nbr_out (may be NULL) receives it,
 * k*(n1+n2) indices. */
int hor_mpnn_set_geometry(hor_mpnn* net, int n1, int n2, const double* x,
                          const double* nrml, const double* ctr, const double* size,
                          int batch_max, int* nbr_out);
/* HOR_OK if the geometry has n/3 points. */
int hor_mpnn_check_size(hor_mpnn* net, int n);
/* z = M(r) for batch vectors of 3*(n1+n2) values each (device pointers). */
int hor_mpnn_forward_dev(hor_mpnn* net, int batch, const double* r_dev, double* z_dev);
/* z = M(r) on ctx's stream without synchronizing (the solver's call). */
int hor_mpnn_apply_dev(hor_mpnn* net, hor_ctx* ctx, const double* r_dev, double* z_dev);
/* Forward on r, then the loss mean_b |M(r_b) - v_b|^2/|v_b|^2 and its gradient
 * with respect to the parameters (kept on the device). */
int hor_mpnn_loss_grad_dev(hor_mpnn* net, int batch, const double* r_dev,
                           const double* v_dev, double* loss);
/* One Adam step with the last gradient. */
int hor_mpnn_adam_step(hor_mpnn* net, float lr, float beta1, float beta2, float eps);

/* Flexible GMRES(m) with the right preconditioner precond (NULL: none, the
 * plain method).  The solve always ends on the FP64 true residual: *relres is
 * |b - A x|/|b| with the FP64 operator.  When log_z_dev / log_az_dev are given
 * (device buffers of log_max*n values), the first log_max Arnoldi pairs
 * (z_j, A z_j) are copied there and *n_logged is their count. */
int hor_solver_fgmres(hor_solver* solver, const double* b, double* x, int m,
                      double tol, int maxit, hor_mpnn* precond, int* iters,
                      double* relres, double* hist, double* log_z_dev,
                      double* log_az_dev, int log_max, int* n_logged);

/* How hor_solver_fgmres uses its network (no counterpart in the reference).
 *   force_plain: nonzero ignores the network: plain GMRES (the host-forced
 *     fallback).
 *   gate: nonzero (the default) makes the first Arnoldi step of every restart
 *     cycle plain and lets its pair (v0, w = A v0) decide the rest of the
 *     cycle: with the scale-free misfit e(X)^2 = 1 - (X(w).v0)^2/(|X(w)|^2
 *     |v0|^2) of an approximate inverse X, the network is used for the
 *     cycle's remaining steps only if e(M) < gate_ratio * e(I).  A cycle whose
 *     gate rejects the network is exactly plain GMRES.  While the network is
 *     in use, a step that reduces the residual estimate by less than the
 *     plain first step did (by the factor e(I)) ends its use for the rest of
 *     the cycle.  With gate = 0 every step uses the network.
 * Every network output is screened: a non-finite one is replaced by the plain
 * step and the network is not used again in that solve.  Defaults
 * {0, 1, 0.5}. */
typedef struct {
  int force_plain;
  int gate;
  double gate_ratio;
} hor_precond_policy;
int hor_solver_set_precond_policy(hor_solver* solver, const hor_precond_policy* policy);

/* Record of the last hor_solver_gmres / hor_solver_fgmres on this solver. */
typedef struct {
  int iterations;
  int learned_steps;        /* Arnoldi steps with z = M(v) */
  int plain_steps;          /* Arnoldi steps with z = v (fgmres) */
  int nonfinite_outputs;    /* network outputs rejected by the finite screen */
  int cycles_gated_on;      /* restart cycles whose gate accepted the network */
  int cycles_gated_off;     /* ... rejected it */
  int monitor_stops;        /* cycles where the progress monitor ended its use */
  double gate_misfit_identity;  /* e(I) and e(M) of the last gate test (-1: none) */
  double gate_misfit_network;
  double relres;            /* the returned relative residual */
} hor_fgmres_info;
int hor_solver_fgmres_info(hor_solver* solver, hor_fgmres_info* info);

/* ------------------------------------------------------------------------
 * Device memory helpers (for callers who keep data resident).
 * ------------------------------------------------------------------------ */
int hor_dev_alloc(hor_ctx* ctx, size_t bytes, void** ptr);
int hor_dev_free(hor_ctx* ctx, void* ptr);
int hor_dev_upload(hor_ctx* ctx, void* dst_dev, const void* src, size_t bytes);
int hor_dev_download(hor_ctx* ctx, void* dst, const void* src_dev, size_t bytes);

/* ------------------------------------------------------------------------
 * Host geometry (input producers; run once per geometry, on the CPU).
 * ------------------------------------------------------------------------ */
/* Level-set body: ellipsoid with semi-axes abc, centre ctr, un-rotation matrix
 * R (row-major 3x3, x_body = R (x - ctr); identity if NULL), as the
 * reference's phi/Dphi (highreg_stokes_2surf/utilities.cpp:16-158). */
typedef struct {
  double abc[3];
  double ctr[3];
  double R[9];
  int rotated;
} hor_body;

int hor_body_init(hor_body* body, const double* abc, const double* ctr, const double* R);

/* Generate_Surface (surface.cpp:13-71): intersections of the grid lines
 * (multiples of h through the body centre) with the body, partition-of-unity
 * weights with angle pou_deg (70 in the reference).  Returns the count in *n;
 * arrays filled when cap >= *n (x, nrml: 3n; area: n; plane: n). */
int hor_surface_generate(const hor_body* body, double h, double pou_deg, int cap,
                         int* n, double* x, double* nrml, double* area, int* plane);

/* Compute_nearest_data (surface.cpp:450-526) of targets against a surface of
 * the body: nrst (0 when |b|/DEL > 8), x0, n0, b, i123.  The surface force
 * (surfc_force, surface.cpp:530-559) at x0 is returned in f0 when non-NULL,
 * zero for far targets. */
int hor_nearest_data(const hor_body* body, double h, double DEL, int ns,
                     const double* sx, const int* splane, int nt, const double* tx,
                     int* nrst, double* x0, double* n0, double* b, int* i123,
                     double* f0);

/* surfc_force (surface.cpp:530-559): f = 2 H gamma n - grad_s gamma with
 * gamma = 1 + z^2 (z relative to the centre) and H the mean curvature. */
int hor_surface_force(const hor_body* body, int n, const double* x,
                      const double* nrml, double* f);

/* Grid nodes origin + h*(i,j,k), 0 <= i,j,k <= ngrid, with |b| <= band
 * (find_nrst_surfc, extend_to_grid/surface.cpp:303-392): returns the count;
 * fills when cap >= count: ijk (3n), x (3n), b, x0 (3n), n0 (3n). */
int hor_band_targets(const hor_body* body, const double* origin, double h,
                     int ngrid, double band, int ns, const double* sx, int cap,
                     int* count, int* ijk, double* x, double* b, double* x0,
                     double* n0);

/* ------------------------------------------------------------------------
 * GPU geometry: the same functions on the device (one context's stream).
 * Host pointers in and out.  Positions, normals, planes, nearest data and
 * stencils are bit-identical to the host functions above; the quadrature
 * weight may differ in its last bits (the partition-of-unity factor uses the
 * device's exp and acos).
 * ------------------------------------------------------------------------ */
/* hor_surface_generate (Generate_Surface, surface.cpp:13-71) on the GPU. */
int hor_surface_generate_gpu(hor_ctx* ctx, const hor_body* body, double h,
                             double pou_deg, int cap, int* n, double* x,
                             double* nrml, double* area, int* plane);

/* hor_nearest_data (Compute_nearest_data, surface.cpp:450-526) on the GPU:
 * the same nearest quadrature point (first index among equal distances),
 * Newton projection, far sentinel and surface force. */
int hor_nearest_data_gpu(hor_ctx* ctx, const hor_body* body, double h,
                         double DEL, int ns, const double* sx, const int* splane,
                         int nt, const double* tx, int* nrst, double* x0,
                         double* n0, double* b, int* i123, double* f0);

/* 6x6 Monge stencils of targets on a surface with centre ctr
 * (interpolate_DL_density + sixPoint_stencil + sixPoint_coeff,
 * main.cpp:682-757, 916-1046), as the solver builds them: stencil[36 nt]
 * source indices [k1][k2] (0 where no point matches, the reference's unset
 * value; all zero for far targets), coeff[12 nt] the Lagrange weights along
 * i1 then i2. */
int hor_stencil_build(hor_ctx* ctx, const double* ctr, double h, int ns,
                      const double* sx, int nt, const int* nrst,
                      const double* x0, const int* i123, int* stencil,
                      double* coeff);

/* Translating-body densities and exact velocity.
 * hor_traction_translating: single-layer density of a body translating with
 * unit velocity U (for an ellipsoid: F/(4 pi abc) / sqrt(x^2/a^4+...)), which
 * makes SL[f] = U inside and on the surface.
 * hor_spheroid_exact_velocity: exterior velocity of a prolate spheroid
 * (a > b = c) translating with U = (1,0,0) along its axis
 * (extend_to_grid/utilities.cpp:360-411, Chwang & Wu). */
int hor_traction_translating(const hor_body* body, const double* U, int n,
                             const double* x, double* f);
int hor_spheroid_exact_velocity(const hor_body* body, int n, const double* x,
                                double* u);
/* Exact velocity of any ellipsoid body (semi-axes abc, rotation R) translating
 * with velocity U through fluid at rest at infinity (Oberbeck's solution in
 * ellipsoidal integrals, Carlson R_F / R_D): U inside and on the body, the
 * exterior Stokes flow outside.  For a prolate spheroid translating along its
 * axis it is SL_exact (extend_to_grid/utilities.cpp:360-411); its force is
 * that of hor_traction_translating's density.  No reference counterpart for
 * other bodies. */
int hor_translating_exact_velocity(const hor_body* body, const double* U, int n,
                                   const double* x, double* u);
/* p_exact (extend_to_grid/utilities.cpp:294-321): pressure of that spheroid. */
int hor_spheroid_exact_pressure(const hor_body* body, int n, const double* x,
                                double* p);
/* The body's level set phi (utilities.cpp:58-86) at n points (host). */
int hor_level_set(const hor_body* body, int n, const double* x, double* phi);

/* ------------------------------------------------------------------------
 * Grid extension (host, OpenMP; highreg_stokes_extend_to_grid/main.cpp).
 * Grid nodes (i,j,k), 0 <= i,j,k <= N, at (i,j,k) L/N, stored at
 * (i (N+1) + j) (N+1) + k.
 * ------------------------------------------------------------------------ */
/* fftw_DST_Poisson3d_FD (main.cpp:1423-1491): f[n^3] = h^2 times the right-hand
 * side at the interior nodes (row-major) in; the solution of the 4th-order
 * discrete Poisson problem with zero Dirichlet data out (DST-I through the
 * library's own FFT; no FFTW). */
int hor_dst_poisson3d(int n, double* f);

/* solve_Poisson (main.cpp:969-1044): extends u[(N+1)^3], given on the box
 * faces and at the band nodes, to the whole grid in place.
 *   band[(N+1)^3]: 0 outside BAND1, 1 in BAND1 only, 2 in BAND2 (|b| <= BAND2 h);
 *   outside[(N+1)^3]: 1 where phi > 0;
 *   variable: 0 for the pressure (no forcing), 1..3 for a velocity component
 *             (forcing dp/dx_variable from p[(N+1)^3] outside the body and BAND2). */
int hor_grid_solve_poisson(int N, double L, const signed char* band,
                           const unsigned char* outside, const double* p,
                           double* u, int variable);

/* k nearest neighbours of each of n points (host): nbr[i*k + a], sorted by
 * (distance, index), the point itself excluded; the graph of the learned
 * preconditioner. */
int hor_knn_graph(int n, const double* x, int k, int* nbr);

#ifdef __cplusplus
}
#endif
#endif
