// Two-surface second-kind solve on the device (highreg_stokes_2surf/main.cpp:140-677).
//
// The matvec A(u) = u - 2/3 [DL11 + DL21 ; DL12 + DL22](u) is four operator
// plans evaluated straight into the output vector, after the density at each
// target's nearest point is interpolated on its 6x6 Monge stencil.  The
// stencil indices and Lagrange weights depend on geometry only, so they are
// found once, on the device, at set-up (the reference re-scans every
// quadrature point for every target on every matvec, main.cpp:969-1006).
// GMRES keeps its Krylov basis in HBM; only the Hessenberg column and the
// rotations live on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "geom_gpu.h"
#include "plan.h"

using namespace hor;

namespace hor {

// g0 = sum_j b_j sum_i a_i g[st[i][j]] (sixPointInterpolate, main.cpp:759-778).
__global__ void k_interp(int nt, const int* nrst, const int* st, const double* co, const double* g,
                         double* g0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  if (nrst[i] == 0) {
    g0[3 * i] = g0[3 * i + 1] = g0[3 * i + 2] = 0.0;
    return;
  }
  const int* s = st + 36 * (size_t)i;
  const double* a = co + 12 * (size_t)i;
  const double* b = a + 6;
  for (int c = 0; c < 3; c++) {
    double v = 0.0;
    for (int j = 0; j < 6; j++) {
      double f = 0.0;
      for (int k = 0; k < 6; k++) f += a[k] * g[3 * s[6 * k + j] + c];
      v += b[j] * f;
    }
    g0[3 * i + c] = v;
  }
}

// ---- vector kernels for GMRES
constexpr int kRedBlocks = 1024;

__global__ void k_dot_partial(int n, const double* x, const double* y, double* part) {
  __shared__ double sm[kBlock];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i] * y[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_axpy(int n, double a, const double* x, double* y) {  // y += a x
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__global__ void k_scale(int n, double a, const double* x, double* y) {  // y = a x
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ void k_sub(int n, const double* a, const double* b, double* y) {  // y = a - b
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] - b[i];
}

}  // namespace hor

struct hor_solver {
  hor_ctx* ctx = nullptr;
  double h = 0, DEL = 0;
  int method = 0, precision = 0;
  int n1 = 0, n2 = 0, n = 0;
  hor_tree_params tp[2];
  // host copies needed for lazily built SL plans
  std::vector<double> x[2], nrml[2], area[2], x0[2], n0[2], b[2];
  double ctr[2][3];
  hor_plan* on[2] = {nullptr, nullptr};
  hor_plan* off[2] = {nullptr, nullptr};   // off[s]: sources of surface s at the other surface's points
  hor_plan* sl_on[2] = {nullptr, nullptr};
  hor_plan* sl_off[2] = {nullptr, nullptr};
  // mixed precision: FP64 operator for the true residual (built on first use)
  hor_plan* on64[2] = {nullptr, nullptr};
  hor_plan* off64[2] = {nullptr, nullptr};
  DBuf<int> nrst[2], st[2];
  DBuf<double> co[2], g0[2];
  DBuf<double> vin, vout, V, Z, w, part, bvec, xvec, r, gt;
  std::vector<double> part_h;
  // learned-preconditioner policy of hor_solver_fgmres and the last solve's record
  hor_precond_policy policy{0, 1, 0.5};
  hor_fgmres_info info{};
  ~hor_solver() {
    for (int k = 0; k < 2; k++) {
      hor_plan_destroy(on[k]);
      hor_plan_destroy(off[k]);
      hor_plan_destroy(sl_on[k]);
      hor_plan_destroy(sl_off[k]);
      hor_plan_destroy(on64[k]);
      hor_plan_destroy(off64[k]);
    }
  }
};

static int dot_dev(hor_solver* S, const double* x, const double* y, double* out) {
  hipStream_t s = S->ctx->stream;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kBlock), 0, s, S->n, x, y, S->part.p);
  HOR_HIP_TRY(hipGetLastError());
  HOR_HIP_TRY(hipMemcpyAsync(S->part_h.data(), S->part.p, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  double v = 0.0;
  for (int k = 0; k < kRedBlocks; k++) v += S->part_h[k];
  *out = v;
  return HOR_OK;
}

static inline dim3 grid_for(int n) { return dim3((n + kBlock - 1) / kBlock); }

// A(u) on device pointers (main.cpp:541-677) with the given DL plans.
static int matvec_plans(hor_solver* S, hor_plan* const on[2], hor_plan* const off[2], const double* u,
                        double* Au) {
  hipStream_t s = S->ctx->stream;
  const double factor = 2.0 / 3.0;
  const double* u1 = u;
  const double* u2 = u + 3 * (size_t)S->n1;
  double* A1 = Au;
  double* A2 = Au + 3 * (size_t)S->n1;
  // g0 at the points of surface 2 from the density of surface 1, and vice versa
  hipLaunchKernelGGL(k_interp, grid_for(S->n2), dim3(kBlock), 0, s, S->n2, S->nrst[1].p, S->st[1].p,
                     S->co[1].p, u1, S->g0[1].p);
  hipLaunchKernelGGL(k_interp, grid_for(S->n1), dim3(kBlock), 0, s, S->n1, S->nrst[0].p, S->st[0].p,
                     S->co[0].p, u2, S->g0[0].p);
  HOR_HIP_TRY(hipGetLastError());
  HOR_TRY(plan_eval_into(on[0], u1, nullptr, A1, u1, -factor, 0));           // u1 - 2/3 DL11
  HOR_TRY(plan_eval_into(off[0], u1, S->g0[1].p, A2, u2, -factor, 0));       // u2 - 2/3 DL12
  // on[k] before off[k]: a treecode off[k] reads the moments on[k] computes
  HOR_TRY(plan_eval_into(on[1], u2, nullptr, A2, nullptr, -factor, 1));      // - 2/3 DL22
  HOR_TRY(plan_eval_into(off[1], u2, S->g0[0].p, A1, nullptr, -factor, 1));  // - 2/3 DL21
  return HOR_OK;
}

// The solver's operator (its own precision).
static int matvec_dev(hor_solver* S, const double* u, double* Au) {
  return matvec_plans(S, S->on, S->off, u, Au);
}

// FP64 plans of a mixed-precision solver, for the true residual.
static int ensure_fp64_plans(hor_solver* S) {
  if (S->precision != HOR_MIXED || S->on64[0]) return HOR_OK;
  hor_ctx* c = S->ctx;
  for (int k = 0; k < 2; k++) {
    const int o = 1 - k;
    const int nk = (int)S->area[k].size(), no = (int)S->area[o].size();
    HOR_TRY(hor_plan_create(c, HOR_DL_ON, S->method, HOR_FP64, S->DEL, nk, S->x[k].data(), S->nrml[k].data(),
                            S->area[k].data(), nk, S->x[k].data(), nullptr, nullptr, nullptr, S->ctr[k],
                            &S->tp[k], &S->on64[k]));
    HOR_TRY(hor_plan_create(c, HOR_DL_OFF, S->method, HOR_FP64, S->DEL, nk, S->x[k].data(), S->nrml[k].data(),
                            S->area[k].data(), no, S->x[o].data(), S->b[o].data(), S->x0[o].data(),
                            S->n0[o].data(), S->ctr[k], &S->tp[k], &S->off64[k]));
    if (S->method == HOR_TREE) {
      HOR_TRY(tree_plan_prepare_subtraction(S->on64[k]));
      HOR_TRY(tree_plan_prepare_subtraction(S->off64[k]));
      HOR_TRY(tree_plan_share_moments(S->on64[k], S->off64[k], S->x[k].data()));
    }
  }
  return HOR_OK;
}

// The FP64 operator: the solver's own plans, or for a mixed-precision solver
// its second set of FP64 plans.
static int matvec64_dev(hor_solver* S, const double* u, double* Au) {
  if (S->precision != HOR_MIXED) return matvec_plans(S, S->on, S->off, u, Au);
  HOR_TRY(ensure_fp64_plans(S));
  return matvec_plans(S, S->on64, S->off64, u, Au);
}

extern "C" int hor_solver_create(hor_ctx* ctx, double h, double DEL, int method, int precision,
                                 const hor_tree_params* tree1, const hor_tree_params* tree2,
                                 const hor_surface_desc* s1, const hor_surface_desc* s2,
                                 hor_solver** out) {
  if (!ctx || !s1 || !s2 || !out) return set_error(HOR_E_ARG, "hor_solver_create: null argument");
  *out = nullptr;
  if (!(h > 0) || !(DEL > 0)) return set_error(HOR_E_ARG, "hor_solver_create: h and DEL must be > 0");
  const hor_surface_desc* sd[2] = {s1, s2};
  for (int k = 0; k < 2; k++)
    if (sd[k]->n <= 0 || !sd[k]->x || !sd[k]->nrml || !sd[k]->area || !sd[k]->nrst || !sd[k]->x0 ||
        !sd[k]->n0 || !sd[k]->b || !sd[k]->i123)
      return set_error(HOR_E_ARG, "hor_solver_create: surface %d incomplete", k + 1);
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  std::unique_ptr<hor_solver> S(new hor_solver());
  S->ctx = ctx;
  S->h = h;
  S->DEL = DEL;
  S->method = method;
  S->precision = precision;
  S->n1 = s1->n;
  S->n2 = s2->n;
  S->n = 3 * (S->n1 + S->n2);
  S->tp[0] = tree1 ? *tree1 : hor_tree_params{6, 1000, 0.6};
  S->tp[1] = tree2 ? *tree2 : hor_tree_params{6, 1000, 0.6};
  hipStream_t st = ctx->stream;
  for (int k = 0; k < 2; k++) {
    const hor_surface_desc* d = sd[k];
    int n = d->n;
    S->x[k].assign(d->x, d->x + 3 * n);
    S->nrml[k].assign(d->nrml, d->nrml + 3 * n);
    S->area[k].assign(d->area, d->area + n);
    S->x0[k].assign(d->x0, d->x0 + 3 * n);
    S->n0[k].assign(d->n0, d->n0 + 3 * n);
    S->b[k].assign(d->b, d->b + n);
    for (int c = 0; c < 3; c++) S->ctr[k][c] = d->ctr[c];
  }
  for (int k = 0; k < 2; k++) {
    int o = 1 - k;
    const hor_surface_desc* me = sd[k];
    const hor_surface_desc* other = sd[o];
    // on-surface DL of surface k
    HOR_TRY(hor_plan_create(ctx, HOR_DL_ON, method, precision, DEL, me->n, me->x, me->nrml, me->area, me->n,
                            me->x, nullptr, nullptr, nullptr, me->ctr, &S->tp[k], &S->on[k]));
    // DL of surface k at the points of the other surface (whose nearest data is relative to k)
    HOR_TRY(hor_plan_create(ctx, HOR_DL_OFF, method, precision, DEL, me->n, me->x, me->nrml, me->area,
                            other->n, other->x, other->b, other->x0, other->n0, me->ctr, &S->tp[k],
                            &S->off[k]));
    if (method == HOR_TREE) {
      HOR_TRY(tree_plan_prepare_subtraction(S->on[k]));
      HOR_TRY(tree_plan_prepare_subtraction(S->off[k]));
      HOR_TRY(tree_plan_share_moments(S->on[k], S->off[k], me->x));
    }
    // stencils of the points of surface k (their nearest points lie on the other surface)
    HOR_TRY(S->st[k].alloc(36 * (size_t)me->n));
    HOR_TRY(S->co[k].alloc(12 * (size_t)me->n));
    HOR_TRY(stencil_build(ctx, other->ctr, h, other->n, other->x, me->n, me->nrst, me->x0, me->i123, S->st[k].p,
                          S->co[k].p));
    HOR_TRY(S->nrst[k].upload(me->nrst, me->n, st));
    HOR_TRY(S->g0[k].alloc(3 * (size_t)me->n));
  }
  HOR_TRY(ensure_fp64_plans(S.get()));
  HOR_TRY(S->vin.alloc(S->n));
  HOR_TRY(S->vout.alloc(S->n));
  HOR_TRY(S->part.alloc(kRedBlocks));
  S->part_h.assign(kRedBlocks, 0.0);
  HOR_HIP_TRY(hipStreamSynchronize(st));
  *out = S.release();
  return HOR_OK;
}

extern "C" int hor_solver_destroy(hor_solver* S) {
  if (!S) return HOR_OK;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  delete S;
  return HOR_OK;
}

// Work of one matvec, summed over its four DL plans (hor_plan_stats).
extern "C" int hor_solver_stats(hor_solver* S, double* direct_pairs, double* near_pairs, double* proxy_pairs,
                                double* moment_terms) {
  if (!S) return set_error(HOR_E_ARG, "hor_solver_stats: null");
  double acc[4] = {0, 0, 0, 0};
  hor_plan* plans[4] = {S->on[0], S->on[1], S->off[0], S->off[1]};
  for (hor_plan* P : plans) {
    double v[4];
    HOR_TRY(hor_plan_stats(P, &v[0], &v[1], &v[2], &v[3]));
    for (int k = 0; k < 4; k++) acc[k] += v[k];
  }
  if (direct_pairs) *direct_pairs = acc[0];
  if (near_pairs) *near_pairs = acc[1];
  if (proxy_pairs) *proxy_pairs = acc[2];
  if (moment_terms) *moment_terms = acc[3];
  return HOR_OK;
}

// Work of one of the four DL plans: 0, 1 the on-surface plans of surfaces 1
// and 2, 2, 3 the plans of surface 1 and 2 sources at the other surface.
extern "C" int hor_solver_plan_stats(hor_solver* S, int which, double* direct_pairs, double* near_pairs,
                                     double* proxy_pairs, double* moment_terms) {
  if (!S || which < 0 || which > 3) return set_error(HOR_E_ARG, "hor_solver_plan_stats: bad argument");
  hor_plan* plans[4] = {S->on[0], S->on[1], S->off[0], S->off[1]};
  return hor_plan_stats(plans[which], direct_pairs, near_pairs, proxy_pairs, moment_terms);
}

// Kernel time by phase of the last evaluation of the solver's four DL plans
// (the operator of its precision), summed: moments, near sums, far field.
extern "C" int hor_solver_phase_timing(hor_solver* S, double* moments_ms, double* near_ms, double* far_ms) {
  if (!S) return set_error(HOR_E_ARG, "hor_solver_phase_timing: null");
  if (S->method != HOR_TREE) return set_error(HOR_E_STATE, "hor_solver_phase_timing: treecode solvers only");
  HOR_HIP_TRY(hipSetDevice(S->ctx->device));
  HOR_HIP_TRY(hipStreamSynchronize(S->ctx->stream));
  double acc[3] = {0, 0, 0};
  hor_plan* plans[4] = {S->on[0], S->on[1], S->off[0], S->off[1]};
  for (hor_plan* P : plans) {
    double ms[3];
    HOR_TRY(tree_plan_phase_ms(P, ms));
    for (int k = 0; k < 3; k++) acc[k] += ms[k];
  }
  if (moments_ms) *moments_ms = acc[0];
  if (near_ms) *near_ms = acc[1];
  if (far_ms) *far_ms = acc[2];
  return HOR_OK;
}

extern "C" int hor_solver_size(hor_solver* S, int* n) {
  if (!S || !n) return set_error(HOR_E_ARG, "null");
  *n = S->n;
  return HOR_OK;
}

extern "C" int hor_solver_matvec_dev(hor_solver* S, const double* u_dev, double* Au_dev) {
  if (!S || !u_dev || !Au_dev) return set_error(HOR_E_ARG, "hor_solver_matvec_dev: null");
  hor_ctx* c = S->ctx;
  HOR_HIP_TRY(hipSetDevice(c->device));
  HOR_HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  HOR_TRY(matvec_dev(S, u_dev, Au_dev));
  HOR_HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  HOR_HIP_TRY(hipStreamSynchronize(c->stream));
  float t = 0;
  HOR_HIP_TRY(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
  c->last_total_ms = t;
  return HOR_OK;
}

extern "C" int hor_solver_matvec(hor_solver* S, const double* u, double* Au) {
  if (!S || !u || !Au) return set_error(HOR_E_ARG, "hor_solver_matvec: null");
  hipStream_t s = S->ctx->stream;
  HOR_HIP_TRY(hipSetDevice(S->ctx->device));
  HOR_HIP_TRY(hipMemcpyAsync(S->vin.p, u, sizeof(double) * S->n, hipMemcpyHostToDevice, s));
  HOR_TRY(hor_solver_matvec_dev(S, S->vin.p, S->vout.p));
  HOR_HIP_TRY(hipMemcpyAsync(Au, S->vout.p, sizeof(double) * S->n, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  return HOR_OK;
}

extern "C" int hor_solver_interp_g0(hor_solver* S, int k, const double* u, double* g0) {
  if (!S || !u || !g0 || (k != 1 && k != 2)) return set_error(HOR_E_ARG, "hor_solver_interp_g0: bad argument");
  hipStream_t s = S->ctx->stream;
  HOR_HIP_TRY(hipSetDevice(S->ctx->device));
  int me = k - 1, other = 1 - me;
  int nother = other == 0 ? S->n1 : S->n2, nme = me == 0 ? S->n1 : S->n2;
  HOR_HIP_TRY(hipMemcpyAsync(S->vin.p, u, sizeof(double) * 3 * nother, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_interp, grid_for(nme), dim3(kBlock), 0, s, nme, S->nrst[me].p, S->st[me].p, S->co[me].p,
                     S->vin.p, S->g0[me].p);
  HOR_HIP_TRY(hipGetLastError());
  HOR_HIP_TRY(hipMemcpyAsync(g0, S->g0[me].p, sizeof(double) * 3 * nme, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  return HOR_OK;
}

// eval_SL (main.cpp:421-537).
extern "C" int hor_solver_eval_sl(hor_solver* S, const double* f1, const double* f2, const double* f01,
                                  const double* f02, double* sl) {
  if (!S || !f1 || !f2 || !f01 || !f02 || !sl) return set_error(HOR_E_ARG, "hor_solver_eval_sl: null");
  hor_ctx* c = S->ctx;
  HOR_HIP_TRY(hipSetDevice(c->device));
  for (int k = 0; k < 2; k++) {
    int o = 1 - k;
    int nk = (int)S->area[k].size(), no = (int)S->area[o].size();
    if (!S->sl_on[k])
      HOR_TRY(hor_plan_create(c, HOR_SL_ON, S->method, S->precision, S->DEL, nk, S->x[k].data(),
                              S->nrml[k].data(), S->area[k].data(), nk, S->x[k].data(), nullptr, nullptr,
                              nullptr, S->ctr[k], &S->tp[k], &S->sl_on[k]));
    if (!S->sl_off[k])
      HOR_TRY(hor_plan_create(c, HOR_SL_OFF, S->method, S->precision, S->DEL, nk, S->x[k].data(),
                              S->nrml[k].data(), S->area[k].data(), no, S->x[o].data(), S->b[o].data(),
                              S->x0[o].data(), S->n0[o].data(), S->ctr[k], &S->tp[k], &S->sl_off[k]));
  }
  hipStream_t s = c->stream;
  int n1 = S->n1, n2 = S->n2;
  DBuf<double> F, F0;
  HOR_TRY(F.alloc(3 * (size_t)(n1 + n2)));
  HOR_TRY(F0.alloc(3 * (size_t)(n1 + n2)));
  HOR_HIP_TRY(hipMemcpyAsync(F.p, f1, sizeof(double) * 3 * n1, hipMemcpyHostToDevice, s));
  HOR_HIP_TRY(hipMemcpyAsync(F.p + 3 * n1, f2, sizeof(double) * 3 * n2, hipMemcpyHostToDevice, s));
  HOR_HIP_TRY(hipMemcpyAsync(F0.p, f01, sizeof(double) * 3 * n1, hipMemcpyHostToDevice, s));
  HOR_HIP_TRY(hipMemcpyAsync(F0.p + 3 * n1, f02, sizeof(double) * 3 * n2, hipMemcpyHostToDevice, s));
  double* o1 = S->vout.p;
  double* o2 = S->vout.p + 3 * n1;
  HOR_TRY(plan_eval_into(S->sl_on[0], F.p, nullptr, o1, nullptr, 1.0, 0));            // SL11
  HOR_TRY(plan_eval_into(S->sl_off[0], F.p, F0.p + 3 * n1, o2, nullptr, 1.0, 0));     // SL12
  HOR_TRY(plan_eval_into(S->sl_off[1], F.p + 3 * n1, F0.p, o1, nullptr, 1.0, 1));     // + SL21
  HOR_TRY(plan_eval_into(S->sl_on[1], F.p + 3 * n1, nullptr, o2, nullptr, 1.0, 1));   // + SL22
  HOR_HIP_TRY(hipMemcpyAsync(sl, S->vout.p, sizeof(double) * S->n, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  return HOR_OK;
}

// Restarted flexible GMRES(m) with MGS Arnoldi and Givens rotations
// (main.cpp:210-416), device-resident.  Arnoldi applies A_arn to z_j, where
// z_j = v_j (a plain step) or z_j = M(v_j) (a learned step); FGMRES allows the
// choice to change from step to step.  Each cycle restarts from the true
// residual b - A_true x (the reference restarts from b, main.cpp:286).  With
// recheck, a cycle that meets tol by its estimate is confirmed with A_true and
// continues if the true residual is above tol, and a solve that reaches maxit
// also ends on the true residual: *relres is then always the true residual.
// Without it the solve stops on the estimate, as the reference does.
// Optionally logs the pairs (z_j, A z_j) into device buffers.
//
// Learned steps (a network M is given and the policy does not force plain):
//   - gate: the first step of every cycle is plain.  Its pair (v_0, w = A v_0)
//     tells how well M inverts A on this cycle's data: with the misfit of an
//     approximate inverse X taken up to scale, e(X)^2 = 1 - (X(w).v_0)^2 /
//     (|X(w)|^2 |v_0|^2), the rest of the cycle uses M only if
//     e(M) < ratio * e(I).  A cycle whose gate rejects M is plain GMRES.
//   - progress monitor: the plain first step reduces the residual estimate by
//     rho_0 = e(I); a learned step that reduces it by less (est_j / est_{j-1}
//     > rho_0) ends the network's use for the rest of the cycle.
//   - every network output is screened: a non-finite z_j is replaced by v_j
//     and M is not used again in this solve.
struct PrecondOp {
  hor_mpnn* net = nullptr;
  bool gate = true;
  double ratio = 1.0;
};

static int apply_precond(hor_solver* S, const PrecondOp& M, const double* v, double* z);

static bool all_finite_host(const double* a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(a[i])) return false;
  return true;
}

static int gmres_core(hor_solver* S, bool flexible, const PrecondOp& M, bool mixed_arnoldi, bool recheck,
                      const double* b_host, double* x_host, int m, double tol, int maxit, int* iters,
                      double* relres, double* hist, double* log_z, double* log_az, int log_max, int* n_logged) {
  hor_ctx* c = S->ctx;
  hipStream_t s = c->stream;
  HOR_HIP_TRY(hipSetDevice(c->device));
  const int n = S->n;
  S->info = hor_fgmres_info{};
  S->info.gate_misfit_identity = S->info.gate_misfit_network = -1.0;
  if (!all_finite_host(b_host, (size_t)n)) return set_error(HOR_E_NAN, "GMRES: non-finite right-hand side");
  HOR_TRY(S->V.alloc((size_t)n * (m + 1)));
  if (flexible) HOR_TRY(S->Z.alloc((size_t)n * m));
  if (flexible && M.net) HOR_TRY(S->gt.alloc(n));
  HOR_TRY(S->w.alloc(n));
  HOR_TRY(S->bvec.alloc(n));
  HOR_TRY(S->xvec.alloc(n));
  HOR_TRY(S->r.alloc(n));
  HOR_HIP_TRY(hipMemcpyAsync(S->bvec.p, b_host, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HOR_HIP_TRY(hipMemsetAsync(S->xvec.p, 0, sizeof(double) * n, s));
  auto A_arn = [&](const double* u, double* Au) { return mixed_arnoldi ? matvec_dev(S, u, Au) : matvec64_dev(S, u, Au); };
  std::vector<double> H((size_t)(m + 1) * m, 0.0), g(m + 1, 0.0), y(m + 1, 0.0), cs(m, 0.0), sn(m, 0.0);
  auto Hh = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
  auto Vp = [&](int i) { return S->V.p + (size_t)i * n; };
  auto Zp = [&](int i) { return flexible ? S->Z.p + (size_t)i * n : S->V.p + (size_t)i * n; };
  dim3 gr = grid_for(n), bl(kBlock);
  double bnorm2;
  HOR_TRY(dot_dev(S, S->bvec.p, S->bvec.p, &bnorm2));
  const double b_norm = std::sqrt(bnorm2);
  int iter = 0, logged = 0;
  double error = 1.0;
  bool have_x = false;
  bool net_alive = flexible && M.net != nullptr;   // cleared for good by a non-finite output
  while (true) {
    // r = b - A_true x (x = 0 on the first cycle: r = b)
    if (have_x) {
      if (iter >= maxit && !recheck) break;
      HOR_TRY(matvec64_dev(S, S->xvec.p, S->r.p));
      hipLaunchKernelGGL(k_sub, gr, bl, 0, s, n, S->bvec.p, S->r.p, Vp(0));
    } else {
      HOR_HIP_TRY(hipMemcpyAsync(Vp(0), S->bvec.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    double rn2;
    HOR_TRY(dot_dev(S, Vp(0), Vp(0), &rn2));
    const double r_norm = std::sqrt(rn2);
    if (!std::isfinite(r_norm)) return set_error(HOR_E_NAN, "GMRES: non-finite residual after %d iterations", iter);
    error = r_norm;
    if (b_norm > 1.e-8) error /= b_norm;
    if (error <= tol || r_norm == 0.0 || iter >= maxit) break;
    hipLaunchKernelGGL(k_scale, gr, bl, 0, s, n, 1.0 / r_norm, Vp(0), Vp(0));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = r_norm;
    int j = -1;
    double est = error;
    // this cycle's steps use the network: decided by the gate after step 0
    bool use_net = net_alive && !M.gate;
    double rho0 = -1.0;  // residual reduction of the cycle's plain first step
    while ((j < m - 1) && (est > tol) && (iter < maxit)) {
      j = j + 1;
      iter = iter + 1;
      const double est_prev = est;
      bool learned = false;
      if (flexible) {
        if (use_net && net_alive) {
          HOR_TRY(apply_precond(S, M, Vp(j), Zp(j)));
          double zz;
          HOR_TRY(dot_dev(S, Zp(j), Zp(j), &zz));
          if (!std::isfinite(zz)) {  // screen: fall back to the plain step
            S->info.nonfinite_outputs++;
            net_alive = false;
          }
        }
        if (use_net && net_alive) {
          S->info.learned_steps++;
          learned = true;
        } else {
          HOR_HIP_TRY(hipMemcpyAsync(Zp(j), Vp(j), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
          S->info.plain_steps++;
        }
      }
      HOR_TRY(A_arn(Zp(j), Vp(j + 1)));
      if (j == 0 && net_alive && M.gate) {
        // gate on the pair (v_0, w = A v_0), |v_0| = 1
        HOR_TRY(apply_precond(S, M, Vp(1), S->gt.p));
        double wv, ww, tv, tt;
        HOR_TRY(dot_dev(S, Vp(1), Vp(0), &wv));
        HOR_TRY(dot_dev(S, Vp(1), Vp(1), &ww));
        HOR_TRY(dot_dev(S, S->gt.p, Vp(0), &tv));
        HOR_TRY(dot_dev(S, S->gt.p, S->gt.p, &tt));
        if (!std::isfinite(tt) || !std::isfinite(tv)) {
          S->info.nonfinite_outputs++;
          net_alive = false;
        } else {
          const double eI = std::sqrt(std::max(0.0, 1.0 - (ww > 0.0 ? wv * wv / ww : 0.0)));
          const double eM = std::sqrt(std::max(0.0, 1.0 - (tt > 0.0 ? tv * tv / tt : 0.0)));
          S->info.gate_misfit_identity = eI;
          S->info.gate_misfit_network = eM;
          use_net = eM < M.ratio * eI;
          if (use_net) S->info.cycles_gated_on++;
          else S->info.cycles_gated_off++;
        }
      }
      if (log_z && logged < log_max) {
        HOR_HIP_TRY(hipMemcpyAsync(log_z + (size_t)logged * n, Zp(j), sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, s));
        HOR_HIP_TRY(hipMemcpyAsync(log_az + (size_t)logged * n, Vp(j + 1), sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, s));
        logged++;
      }
      for (int i = 0; i <= j; i++) {  // modified Gram-Schmidt
        double t;
        HOR_TRY(dot_dev(S, Vp(i), Vp(j + 1), &t));
        Hh(i, j) = t;
        hipLaunchKernelGGL(k_axpy, gr, bl, 0, s, n, -t, Vp(i), Vp(j + 1));
      }
      double vn2;
      HOR_TRY(dot_dev(S, Vp(j + 1), Vp(j + 1), &vn2));
      double vn = std::sqrt(vn2);
      Hh(j + 1, j) = vn;
      if (vn != 0.0) hipLaunchKernelGGL(k_scale, gr, bl, 0, s, n, 1.0 / vn, Vp(j + 1), Vp(j + 1));
      for (int i = 0; i <= j - 1; i++) {
        double t = Hh(i, j);
        Hh(i, j) = cs[i] * t + sn[i] * Hh(i + 1, j);
        Hh(i + 1, j) = -sn[i] * t + cs[i] * Hh(i + 1, j);
      }
      double hh = std::sqrt(Hh(j, j) * Hh(j, j) + Hh(j + 1, j) * Hh(j + 1, j));
      if (hh == 0.0) hh = 1.e-16;
      cs[j] = Hh(j, j) / hh;
      sn[j] = Hh(j + 1, j) / hh;
      Hh(j, j) = cs[j] * Hh(j, j) + sn[j] * Hh(j + 1, j);
      Hh(j + 1, j) = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      est = std::fabs(g[j + 1]);
      if (b_norm > 1.e-8) est /= b_norm;
      if (!std::isfinite(est)) return set_error(HOR_E_NAN, "GMRES: non-finite Arnoldi step at iteration %d", iter);
      if (hist) hist[iter - 1] = est;
      if (j == 0 && !learned && est_prev > 0.0) rho0 = est / est_prev;
      if (learned && M.gate && rho0 > 0.0 && est > rho0 * est_prev) {  // progress monitor
        use_net = false;
        S->info.monitor_stops++;
      }
    }
    if (j < 0) break;  // no step was taken (not reached: error > tol and iter < maxit here)
    // x += Z y with H y = g (upper triangular)
    for (int i = 0; i <= j; i++) y[i] = g[i];
    y[j] = y[j] / Hh(j, j);
    for (int k = j - 1; k >= 0; k--) {
      double t = y[k];
      for (int i = k + 1; i <= j; i++) t -= Hh(k, i) * y[i];
      y[k] = t / Hh(k, k);
    }
    for (int i = 0; i <= j; i++) hipLaunchKernelGGL(k_axpy, gr, bl, 0, s, n, y[i], Zp(i), S->xvec.p);
    HOR_HIP_TRY(hipGetLastError());
    have_x = true;
    error = est;
    if (!recheck && est <= tol) break;
  }
  HOR_HIP_TRY(hipMemcpyAsync(x_host, S->xvec.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  if (iters) *iters = iter;
  if (relres) *relres = error;
  if (n_logged) *n_logged = logged;
  S->info.iterations = iter;
  S->info.relres = error;
  return HOR_OK;
}

// GMRES(m) of test_Integral_Equation_gmres.  FP64: stops on the residual
// estimate as the reference does.  Mixed precision: Arnoldi with the mixed
// operator, restarts and the final check with the FP64 operator.
extern "C" int hor_solver_gmres(hor_solver* S, const double* b, double* x, int m, double tol, int maxit,
                                int* iters, double* relres, double* hist) {
  if (!S || !b || !x || m < 1 || maxit < 0) return set_error(HOR_E_ARG, "hor_solver_gmres: bad argument");
  const bool mixed = S->precision == HOR_MIXED;
  return gmres_core(S, false, PrecondOp{}, mixed, mixed, b, x, m, tol, maxit, iters, relres, hist, nullptr,
                    nullptr, 0, nullptr);
}

// Flexible GMRES(m) with an optional right preconditioner (a trained
// point-cloud network) under the solver's policy; the final residual is always
// the FP64 true residual.
extern "C" int hor_solver_fgmres(hor_solver* S, const double* b, double* x, int m, double tol, int maxit,
                                 hor_mpnn* precond, int* iters, double* relres, double* hist, double* log_z_dev,
                                 double* log_az_dev, int log_max, int* n_logged) {
  if (!S || !b || !x || m < 1 || maxit < 0 || log_max < 0 || ((log_z_dev == nullptr) != (log_az_dev == nullptr)))
    return set_error(HOR_E_ARG, "hor_solver_fgmres: bad argument");
  PrecondOp M;
  M.net = S->policy.force_plain ? nullptr : precond;
  M.gate = S->policy.gate != 0;
  M.ratio = S->policy.gate_ratio;
  if (precond) HOR_TRY(hor_mpnn_check_size(precond, S->n));
  return gmres_core(S, true, M, S->precision == HOR_MIXED, true, b, x, m, tol, maxit, iters, relres, hist,
                    log_z_dev, log_az_dev, log_max, n_logged);
}

extern "C" int hor_solver_set_precond_policy(hor_solver* S, const hor_precond_policy* p) {
  if (!S || !p) return set_error(HOR_E_ARG, "hor_solver_set_precond_policy: null");
  if (!(p->gate_ratio > 0.0) || !std::isfinite(p->gate_ratio))
    return set_error(HOR_E_ARG, "hor_solver_set_precond_policy: gate_ratio must be finite and > 0");
  S->policy = *p;
  return HOR_OK;
}

extern "C" int hor_solver_fgmres_info(hor_solver* S, hor_fgmres_info* info) {
  if (!S || !info) return set_error(HOR_E_ARG, "hor_solver_fgmres_info: null");
  *info = S->info;
  return HOR_OK;
}

static int apply_precond(hor_solver* S, const PrecondOp& M, const double* v, double* z) {
  if (!M.net) {
    HOR_HIP_TRY(hipMemcpyAsync(z, v, sizeof(double) * S->n, hipMemcpyDeviceToDevice, S->ctx->stream));
    return HOR_OK;
  }
  return hor_mpnn_apply_dev(M.net, S->ctx, v, z);
}
