// Host side of the operator plans: context, geometry-only set-up (sorting,
// regularization coefficients, bounding boxes, treecode lists) and the
// per-evaluation launch sequence.  C ABI in include/hor.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "device.h"
#include "direct.h"
#include "plan.h"

using namespace hor;

// ---------------------------------------------------------------------------
// Context

extern "C" int hor_device_count(int* count) {
  if (!count) return set_error(HOR_E_ARG, "hor_device_count: null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return HOR_OK;
}

extern "C" int hor_ctx_create(int device, hor_ctx** out) {
  if (!out) return set_error(HOR_E_ARG, "hor_ctx_create: null");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return set_error(HOR_E_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= n) return set_error(HOR_E_ARG, "device %d out of range (%d devices)", device, n);
  hipDeviceProp_t prop;
  HOR_HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(HOR_E_NODEVICE, "device %d is %s; this build targets gfx950 only", device, prop.gcnArchName);
  HOR_HIP_TRY(hipSetDevice(device));
  std::unique_ptr<hor_ctx> c(new hor_ctx());
  c->device = device;
  HOR_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int k = 0; k < 4; k++) HOR_HIP_TRY(hipEventCreate(&c->ev[k]));
  *out = c.release();
  return HOR_OK;
}

extern "C" int hor_ctx_destroy(hor_ctx* c) {
  if (!c) return HOR_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 4; k++)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return HOR_OK;
}

extern "C" int hor_ctx_last_timing(hor_ctx* c, double* total_ms, double* kernel_ms) {
  if (!c) return set_error(HOR_E_ARG, "null ctx");
  if (total_ms) *total_ms = c->last_total_ms;
  if (kernel_ms) *kernel_ms = c->last_kernel_ms;
  return HOR_OK;
}

extern "C" int hor_dev_alloc(hor_ctx* c, size_t bytes, void** ptr) {
  if (!c || !ptr) return set_error(HOR_E_ARG, "hor_dev_alloc: null");
  HOR_HIP_TRY(hipSetDevice(c->device));
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return set_error(HOR_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return HOR_OK;
}

extern "C" int hor_dev_free(hor_ctx* c, void* ptr) {
  if (!c) return set_error(HOR_E_ARG, "hor_dev_free: null");
  HOR_HIP_TRY(hipSetDevice(c->device));
  if (ptr) HOR_HIP_TRY(hipFree(ptr));
  return HOR_OK;
}

extern "C" int hor_dev_upload(hor_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return set_error(HOR_E_ARG, "hor_dev_upload: null");
  HOR_HIP_TRY(hipSetDevice(c->device));
  HOR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HOR_HIP_TRY(hipStreamSynchronize(c->stream));
  return HOR_OK;
}

extern "C" int hor_dev_download(hor_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return set_error(HOR_E_ARG, "hor_dev_download: null");
  HOR_HIP_TRY(hipSetDevice(c->device));
  HOR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HOR_HIP_TRY(hipStreamSynchronize(c->stream));
  return HOR_OK;
}

// ---------------------------------------------------------------------------
// Plan set-up

namespace hor {

static inline int round_up(int n, int m) { return ((n + m - 1) / m) * m; }

static bool all_finite(const double* a, size_t n) {
  if (!a) return true;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(a[i])) return false;
  return true;
}

// Regularization class of a target: the reference's target-level test
// (stokeslets.cpp:40 for SL, stresslets.cpp:41 for DL; on the surface the
// regularized kernel is always used).
bool target_is_near(const hor_plan* P, double b) {
  int kind = P->kind;
  double DEL = P->DEL;
  if (kind == HOR_SL_ON || kind == HOR_DL_ON || P->all_near) return true;
  if ((kind == HOR_SL_OFF || kind == HOR_P_OFF || kind == HOR_SLDL_OFF) && P->method == HOR_DIRECT)
    return !(std::fabs(b / DEL) > 8.0);  // the same class as the DL_OFF rule below
  return !(std::fabs(b) / DEL > 8.0);  // stresslets.cpp:41, KITC.cpp:856, 970
}

// Per-target regularization coefficients c1, c2, c3 (see device.h; the sum
// kernels expand them with poly_coefs).  Zero for a target beyond 8 delta and
// for the on-surface stresslet, whose kernel comes from its own table.
void target_coefs(int kind, bool nearp, double b, double DEL, double c[3]) {
  c[0] = c[1] = c[2] = 0.0;
  if (!nearp || kind == HOR_DL_ON) return;
  if (kind == HOR_SL_ON) {
    const double rp = 1.7724538509055160273;
    c[0] = 22.0 / 5.0 / rp;
    c[1] = 8.0 / 5.0 / rp;
    c[2] = 2.0 / 15.0 / rp;
  } else {
    reg_coeff_7ord(b / DEL, c);
  }
}

// Sets up the target side shared by the direct and treecode plans: sorted
// order, SoA arrays (padded to ntp with copies of the last target), classes,
// coefficients, wave boxes.
int setup_targets(hor_plan* P, const double* tx, const double* tb, const double* tx0,
                  const double* tn0, const double* sx_orig) {
  hipStream_t s = P->ctx->stream;
  int nt = P->nt, kind = P->kind;
  std::vector<int> cls(nt);
  for (int i = 0; i < nt; i++) cls[i] = target_is_near(P, tb ? tb[i] : 0.0) ? 1 : 0;
  P->tperm_h = spatial_order(nt, tx, cls.data());
  int ntp = std::max(kBlock, round_up(nt, kBlock));
  P->ntp = ntp;
  std::vector<double> X(3 * (size_t)ntp), C(3 * (size_t)ntp, 0.0), X0, N0, Bv, XI;
  std::vector<int> NEAR(ntp);
  bool off_dl = (kind == HOR_DL_OFF || kind == HOR_SLDL_OFF);
  if (off_dl) {
    X0.assign(3 * (size_t)ntp, 0.0);
    N0.assign(3 * (size_t)ntp, 0.0);
    Bv.assign(ntp, 0.0);
  }
  XI.assign(ntp, 0.0);
  for (int i = 0; i < ntp; i++) {
    int o = P->tperm_h[std::min(i, nt - 1)];
    bool nearp = cls[o] != 0;
    NEAR[i] = nearp ? 1 : 0;
    for (int k = 0; k < 3; k++) X[(size_t)k * ntp + i] = tx[3 * o + k];
    double b = tb ? tb[o] : 0.0;
    double c[3];
    target_coefs(kind, nearp, b, P->DEL, c);
    for (int k = 0; k < 3; k++) C[(size_t)k * ntp + i] = c[k];
    if (off_dl) {
      for (int k = 0; k < 3; k++) {
        X0[(size_t)k * ntp + i] = tx0[3 * o + k];
        N0[(size_t)k * ntp + i] = tn0[3 * o + k];
      }
      Bv[i] = b;
    }
    if (i >= nt) continue;
    // jump coefficient of the double layer
    if (kind == HOR_DL_ON) XI[i] = 0.5;                                   // stresslets.cpp:276-278
    if (kind == HOR_DL_OFF || kind == HOR_SLDL_OFF) {
      if (P->method == HOR_DIRECT)
        XI[i] = (nearp && (b < 0.0 || std::fabs(b) < 1.0e-14)) ? 1.0 : 0.0;  // stresslets.cpp:174
      else
        XI[i] = (std::fabs(b) < 1.0e-12) ? 0.5 : ((b < 0.0) ? 1.0 : 0.0);   // KITC.cpp:1805-1811
    }
    if (kind == HOR_P_OFF) XI[i] = (b < -1.e-12) ? -1.0 : 0.0;              // stokes.cpp:198-200
  }
  // wave boxes
  int nw = ntp / kWave;
  std::vector<double> WB(8 * (size_t)nw);
  for (int w = 0; w < nw; w++) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    bool all_far = (kind == HOR_SL_OFF || kind == HOR_DL_OFF || kind == HOR_P_OFF || kind == HOR_SLDL_OFF);
    for (int i = w * kWave; i < (w + 1) * kWave; i++) {
      for (int k = 0; k < 3; k++) {
        double v = X[(size_t)k * ntp + i];
        lo[k] = std::min(lo[k], v);
        hi[k] = std::max(hi[k], v);
      }
      if (NEAR[i]) all_far = false;
    }
    for (int k = 0; k < 3; k++) { WB[8 * w + k] = lo[k]; WB[8 * w + 3 + k] = hi[k]; }
    WB[8 * w + 6] = all_far ? 1.0 : 0.0;
    WB[8 * w + 7] = 0.0;
  }
  HOR_TRY(P->tperm.upload(P->tperm_h, s));
  HOR_TRY(P->tx.upload(X, s));
  HOR_TRY(P->tcoef.upload(C, s));
  HOR_TRY(P->tnear.upload(NEAR, s));
  HOR_TRY(P->wave_box.upload(WB, s));
  HOR_TRY(P->xi.upload(XI, s));
  if (off_dl) {
    HOR_TRY(P->tx0.upload(X0, s));
    HOR_TRY(P->tn0.upload(N0, s));
    HOR_TRY(P->tb.upload(Bv, s));
  }
  if (kind == HOR_SL_OFF || kind == HOR_DL_OFF || kind == HOR_P_OFF || kind == HOR_SLDL_OFF)
    HOR_TRY(P->n0_orig.upload(tn0, 3 * (size_t)nt, s));
  const size_t per = kind == HOR_SLDL_OFF ? 6 : 3;   // d0 and result values per target
  HOR_TRY(P->tsub.alloc((kind == HOR_P_OFF || kind == HOR_SLDL_OFF ? 4 : 3) * (size_t)ntp));
  HOR_TRY(P->tsub2.alloc(3 * (size_t)ntp));
  HOR_TRY(P->in_d0.alloc(per * (size_t)std::max(nt, 1)));
  HOR_TRY(P->out_buf.alloc(per * (size_t)std::max(nt, 1)));
  (void)sx_orig;
  return HOR_OK;
}

// Sorted source arrays (SoA, stride nsp) and their 64-subtile boxes.
int setup_sources(hor_plan* P, const std::vector<int>& perm, const double* sx, const double* sn,
                  const double* sa) {
  hipStream_t s = P->ctx->stream;
  int ns = P->ns;
  int nsp = std::max(kWave, round_up(ns, kWave));
  P->nsp = nsp;
  P->sperm_h = perm;
  if (P->kind == HOR_DL_ON) HOR_TRY(P->dlon_tab.upload(kDlOnTable, kDlOnTabLen, s));
  else HOR_TRY(P->ae_tab.upload(kErfExpTable, kAETabLen, s));
  std::vector<double> X(3 * (size_t)nsp, 0.0), N(3 * (size_t)nsp, 0.0), A((size_t)nsp, 0.0);
  for (int j = 0; j < ns; j++) {
    int o = perm[j];
    for (int k = 0; k < 3; k++) {
      X[(size_t)k * nsp + j] = sx[3 * o + k];
      N[(size_t)k * nsp + j] = sn[3 * o + k];
    }
    A[j] = sa[o];
  }
  int nsub = (ns + kWave - 1) / kWave;
  std::vector<double> SB(6 * (size_t)std::max(nsub, 1), 0.0);
  for (int b = 0; b < nsub; b++) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int j = b * kWave; j < std::min(ns, (b + 1) * kWave); j++)
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], X[(size_t)k * nsp + j]);
        hi[k] = std::max(hi[k], X[(size_t)k * nsp + j]);
      }
    for (int k = 0; k < 3; k++) { SB[6 * b + k] = lo[k]; SB[6 * b + 3 + k] = hi[k]; }
  }
  HOR_TRY(P->sperm.upload(perm, s));
  HOR_TRY(P->sx.upload(X, s));
  HOR_TRY(P->sn.upload(N, s));
  HOR_TRY(P->sa.upload(A, s));
  HOR_TRY(P->sub_box.upload(SB, s));
  HOR_TRY(P->sd.alloc(3 * (size_t)nsp));
  if (P->kind == HOR_SLDL_OFF) {
    HOR_TRY(P->recv.alloc(nsp));
    if (P->precision == HOR_MIXED) HOR_TRY(P->recv32.alloc(nsp));
  } else if (P->kind == HOR_P_OFF) {
    HOR_TRY(P->recp.alloc(nsp));
  } else {
    HOR_TRY(P->rec.alloc(nsp));
    if (P->precision == HOR_MIXED) HOR_TRY(P->rec32.alloc(nsp));
  }
  HOR_TRY(P->sn_orig.upload(sn, 3 * (size_t)ns, s));
  HOR_TRY(P->in_dens.alloc((P->kind == HOR_SLDL_OFF ? 6 : 3) * (size_t)ns));
  return HOR_OK;
}

// Fills the direct-sum argument block for the plan's current buffers.
DirectArgs direct_args(hor_plan* P) {
  DirectArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ns = P->ns; A.nsp = P->nsp; A.nt = P->nt; A.ntp = P->ntp;
  A.nsub = P->nsub; A.subs_per_chunk = P->subs_per_chunk;
  A.DEL = P->DEL;
  double fd = 8.0 * P->DEL * (1.0 + 1e-10);
  A.far_d2 = fd * fd;
  A.near_r2 = 64.0 * P->DEL * P->DEL * (1.0 + 1e-12);
  A.near64_r2 = 16.0 * P->DEL * P->DEL;
  A.sx = P->sx.p; A.sn = P->sn.p; A.sa = P->sa.p; A.sd = P->sd.p; A.sub_box = P->sub_box.p;
  A.rec = P->rec.p; A.recv = P->recv.p; A.recp = P->recp.p;
  A.rec32 = P->precision == HOR_MIXED ? P->rec32.p : nullptr;
  A.recv32 = P->precision == HOR_MIXED ? P->recv32.p : nullptr;
  A.tx = P->tx.p; A.tcoef = P->tcoef.p; A.tsub = P->tsub.p;
  A.tx0 = P->tx0.p; A.tn0 = P->tn0.p; A.tb = P->tb.p;
  A.tnear = P->tnear.p; A.wave_box = P->wave_box.p; A.partial = P->partial.p;
  A.dlon_tab = P->dlon_tab.p;
  A.ae_tab = P->ae_tab.p;
  return A;
}

double op_scale(int kind) {
  const double denom = 1.0 / (8.0 * kPi);                      // utilities.h:11
  if (kind == HOR_P_OFF) return -2.0 * denom;                   // -1/(4 pi), stokes.cpp:196
  return (kind == HOR_SL_ON || kind == HOR_SL_OFF) ? denom : -6.0 * denom;
}

// Per-target subtraction data for an evaluation (device pointers).
int prepare_targets(hor_plan* P, const double* dens_dev, const double* d0_dev) {
  hipStream_t s = P->ctx->stream;
  int mode = (P->kind == HOR_SL_OFF) ? 0 : (P->kind == HOR_SL_ON) ? 1 : (P->kind == HOR_DL_OFF) ? 2 : 3;
  if (P->kind == HOR_P_OFF) mode = 5;
  if (P->kind == HOR_SLDL_OFF) mode = 6;
  if (P->on_at) mode = 4;  // on-surface stresslet at given targets: g0 = d0
  HOR_TRY(launch_target_sub(mode, P->nt, P->ntp, P->tperm.p, P->tnear.p, d0_dev, P->n0_orig.p, dens_dev,
                            P->sn_orig.p, P->tsub.p, s));
  return HOR_OK;
}

}  // namespace hor

static int plan_create_direct(hor_plan* P, const double* sx, const double* sn, const double* sa,
                              const double* tx, const double* tb, const double* tx0,
                              const double* tn0) {
  HOR_TRY(setup_sources(P, spatial_order(P->ns, sx, nullptr), sx, sn, sa));
  HOR_TRY(setup_targets(P, tx, tb, tx0, tn0, sx));
  P->nsub = (P->ns + kWave - 1) / kWave;
  // Source chunks (grid.y): enough workgroups for ~40 rounds of the chip's
  // 2048 resident ones, so the last round's partial fill costs little (one
  // chunk at 1.3M targets leaves 2.5 rounds), and at least 8 subtiles (512
  // sources) per chunk, so the partial-sum traffic stays small against the
  // pair work.  The partial buffer is then <= ~82k * 256 * 6 doubles.
  constexpr int kWantBlocks = 40 * 2048, kMinSubsPerChunk = 8, kMaxChunks = 4096;
  int nblk = P->ntp / kBlock;
  int want = std::max(1, (kWantBlocks + nblk - 1) / nblk);
  int nchunk = std::min({want, std::max(1, P->nsub / kMinSubsPerChunk), kMaxChunks});
  P->subs_per_chunk = (P->nsub + nchunk - 1) / nchunk;
  P->nchunk = (P->nsub + P->subs_per_chunk - 1) / P->subs_per_chunk;
  HOR_TRY(P->partial.alloc((size_t)P->nchunk * (P->kind == HOR_SLDL_OFF ? 6 : 3) * P->ntp));
  P->direct_pairs = (double)P->ns * (double)P->nt;
  return HOR_OK;
}

static int plan_eval_direct(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev,
                            const double* base, double coef, int accumulate) {
  hor_ctx* c = P->ctx;
  hipStream_t s = c->stream;
  if (P->kind == HOR_SLDL_OFF) {
    if (base || coef != 1.0 || accumulate)
      return set_error(HOR_E_ARG, "the fused Stokeslet + stresslet sum writes its result only");
    const bool mixed = P->precision == HOR_MIXED;
    HOR_TRY(launch_gather_sldl(dens_dev, P->sperm.p, P->ns, P->nsp, P->sx.p, P->sn.p, P->sa.p, P->recv.p,
                               mixed ? P->recv32.p : nullptr, P->sub_box.p, s));
    HOR_TRY(prepare_targets(P, dens_dev, d0_dev));
    DirectArgs A = direct_args(P);
    HOR_HIP_TRY(hipEventRecord(c->ev[2], s));
    HOR_TRY(launch_direct(A, P->kind, mixed, P->nchunk, s));
    HOR_HIP_TRY(hipEventRecord(c->ev[3], s));
    return launch_sldl_finalize(P->partial.p, P->nchunk, P->nt, P->ntp, op_scale(HOR_SL_OFF), op_scale(HOR_DL_OFF),
                                P->xi.p, P->tsub.p + P->ntp, P->tperm.p, out_dev, s);
  }
  if (P->kind == HOR_P_OFF)
    HOR_TRY(launch_gather_p(dens_dev, P->sperm.p, P->ns, P->nsp, P->sx.p, P->sn.p, P->sa.p, P->sub_box.p, P->recp.p, s));
  else
    HOR_TRY(launch_gather3(dens_dev, P->sperm.p, P->ns, P->nsp, P->sx.p, P->sn.p, P->sa.p, P->sd.p, P->rec.p,
                           P->precision == HOR_MIXED ? P->rec32.p : nullptr, P->sub_box.p,
                           P->kind == HOR_SL_ON || P->kind == HOR_SL_OFF, s));
  HOR_TRY(prepare_targets(P, dens_dev, d0_dev));
  DirectArgs A = direct_args(P);
  HOR_HIP_TRY(hipEventRecord(c->ev[2], s));
  HOR_TRY(launch_direct(A, P->kind, P->precision == HOR_MIXED, P->nchunk, s));
  HOR_HIP_TRY(hipEventRecord(c->ev[3], s));
  if (P->kind == HOR_P_OFF)
    return launch_pressure_finalize(P->partial.p, P->nchunk, P->nt, P->ntp, op_scale(P->kind), P->xi.p,
                                    P->tsub.p, P->tperm.p, out_dev, base, coef, accumulate, s);
  bool dl = (P->kind == HOR_DL_ON || P->kind == HOR_DL_OFF);
  HOR_TRY(launch_finalize(P->partial.p, P->nchunk, P->nt, P->ntp, op_scale(P->kind), dl ? P->xi.p : nullptr,
                          dl ? P->tsub.p : nullptr, P->tperm.p, out_dev, base, coef, accumulate, s));
  return HOR_OK;
}

// Evaluates the plan into out_dev with the combination rule of the finalize
// kernel (used by the solver to form u - 2/3 DL without extra passes).
int hor::plan_eval_into(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev,
                        const double* base, double coef, int accumulate) {
  HOR_HIP_TRY(hipSetDevice(P->ctx->device));
  if (P->method == HOR_TREE) return tree_plan_eval(P, dens_dev, d0_dev, out_dev, base, coef, accumulate);
  return plan_eval_direct(P, dens_dev, d0_dev, out_dev, base, coef, accumulate);
}

extern "C" int hor_plan_create(hor_ctx* ctx, int kind, int method, int precision, double DEL, int ns,
                               const double* sx, const double* sn, const double* sa, int nt,
                               const double* tx, const double* tb, const double* tx0,
                               const double* tn0, const double* ctr, const hor_tree_params* tree,
                               hor_plan** plan) {
  if (!ctx || !plan) return set_error(HOR_E_ARG, "hor_plan_create: null ctx/plan");
  *plan = nullptr;
  const int flags = kind & 0xF0;
  kind &= 0x0F;
  if (kind < HOR_SL_ON || kind > HOR_SLDL_OFF) return set_error(HOR_E_ARG, "bad kind %d", kind);
  if (kind == HOR_P_OFF && (flags || method != HOR_DIRECT || precision != HOR_FP64))
    return set_error(HOR_E_ARG, "the pressure operator is a direct FP64 sum without kind flags");
  if (kind == HOR_SLDL_OFF && (flags || method != HOR_DIRECT))
    return set_error(HOR_E_ARG, "the fused Stokeslet + stresslet operator is a direct sum without kind flags");
  if ((flags & ~(HOR_ALL_NEAR | HOR_ON_AT)) || ((flags & HOR_ALL_NEAR) && kind != HOR_SL_OFF && kind != HOR_DL_OFF) ||
      ((flags & HOR_ON_AT) && kind != HOR_DL_ON))
    return set_error(HOR_E_ARG, "bad kind flags 0x%x for kind %d", flags, kind);
  if (flags && method != HOR_DIRECT) return set_error(HOR_E_ARG, "kind flags apply to direct sums only");
  if (method != HOR_DIRECT && method != HOR_TREE) return set_error(HOR_E_ARG, "bad method %d", method);
  if (precision != HOR_FP64 && precision != HOR_MIXED) return set_error(HOR_E_ARG, "bad precision %d", precision);
  if (!(DEL > 0) || !std::isfinite(DEL)) return set_error(HOR_E_ARG, "DEL must be finite and > 0");
  if (ns <= 0 || nt <= 0) return set_error(HOR_E_ARG, "need ns > 0 and nt > 0 (ns=%d nt=%d)", ns, nt);
  if (!sx || !sn || !sa || !tx) return set_error(HOR_E_ARG, "null source/target array");
  bool on = (kind == HOR_SL_ON || (kind == HOR_DL_ON && !(flags & HOR_ON_AT)));
  if (on && nt > ns) return set_error(HOR_E_ARG, "on-surface operator needs nt <= ns (target i is source i)");
  if (!on && (!tb || !tn0)) return set_error(HOR_E_ARG, "off-surface operator needs b and nrst normal");
  if ((kind == HOR_DL_OFF || kind == HOR_SLDL_OFF) && !tx0)
    return set_error(HOR_E_ARG, "DL off-surface operator needs nrst point");
  if (method == HOR_TREE && !ctr) return set_error(HOR_E_ARG, "treecode needs the root-box centre");
  if (!all_finite(sx, 3 * (size_t)ns) || !all_finite(sn, 3 * (size_t)ns) || !all_finite(sa, ns) ||
      !all_finite(tx, 3 * (size_t)nt) || !all_finite(tb, tb ? nt : 0) ||
      !all_finite(tx0, tx0 ? 3 * (size_t)nt : 0) || !all_finite(tn0, tn0 ? 3 * (size_t)nt : 0))
    return set_error(HOR_E_NAN, "non-finite geometry input");
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  std::unique_ptr<hor_plan> P(new hor_plan());
  P->ctx = ctx;
  P->kind = kind;
  P->all_near = (flags & HOR_ALL_NEAR) != 0;
  P->on_at = (flags & HOR_ON_AT) != 0;
  P->method = method;
  P->precision = precision;
  P->DEL = DEL;
  P->ns = ns;
  P->nt = nt;
  if (method == HOR_DIRECT) {
    HOR_TRY(plan_create_direct(P.get(), sx, sn, sa, tx, tb, tx0, tn0));
  } else {
    hor_tree_params tp = tree ? *tree : hor_tree_params{6, 1000, 0.6};
    if (tp.P < 1 || tp.P > 12 || tp.N0 < 1 || !(tp.theta > 0))
      return set_error(HOR_E_ARG, "bad tree params P=%d N0=%d theta=%g", tp.P, tp.N0, tp.theta);
    HOR_TRY(tree_plan_create(P.get(), sx, sn, sa, tx, tb, tx0, tn0, ctr, tp));
  }
  HOR_HIP_TRY(hipStreamSynchronize(ctx->stream));
  *plan = P.release();
  return HOR_OK;
}

extern "C" int hor_plan_destroy(hor_plan* P) {
  if (!P) return HOR_OK;
  (void)hipSetDevice(P->ctx->device);
  (void)hipStreamSynchronize(P->ctx->stream);
  delete P;
  return HOR_OK;
}

extern "C" int hor_plan_eval_dev(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev) {
  if (!P || !dens_dev || !out_dev) return set_error(HOR_E_ARG, "hor_plan_eval_dev: null");
  bool on = (P->kind == HOR_SL_ON || P->kind == HOR_DL_ON) && !P->on_at;
  if (!on && !d0_dev) return set_error(HOR_E_ARG, "off-surface operator needs d0");
  hor_ctx* c = P->ctx;
  HOR_HIP_TRY(hipSetDevice(c->device));
  HOR_HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  HOR_TRY(plan_eval_into(P, dens_dev, d0_dev, out_dev, nullptr, 1.0, 0));
  HOR_HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  HOR_HIP_TRY(hipStreamSynchronize(c->stream));
  float t0 = 0, t1 = 0;
  HOR_HIP_TRY(hipEventElapsedTime(&t0, c->ev[0], c->ev[1]));
  HOR_HIP_TRY(hipEventElapsedTime(&t1, c->ev[2], c->ev[3]));
  c->last_total_ms = t0;
  c->last_kernel_ms = t1;
  return HOR_OK;
}

extern "C" int hor_plan_eval(hor_plan* P, const double* dens, const double* d0, double* out) {
  if (!P || !dens || !out) return set_error(HOR_E_ARG, "hor_plan_eval: null");
  bool on = (P->kind == HOR_SL_ON || P->kind == HOR_DL_ON) && !P->on_at;
  if (!on && !d0) return set_error(HOR_E_ARG, "off-surface operator needs d0 (nrst_f / nrst_g)");
  const size_t per = P->kind == HOR_SLDL_OFF ? 6 : 3;   // values per source / target
  if (!all_finite(dens, per * (size_t)P->ns) || (!on && !all_finite(d0, per * (size_t)P->nt)))
    return set_error(HOR_E_NAN, "non-finite density");
  hipStream_t s = P->ctx->stream;
  HOR_HIP_TRY(hipSetDevice(P->ctx->device));
  HOR_HIP_TRY(hipMemcpyAsync(P->in_dens.p, dens, per * sizeof(double) * P->ns, hipMemcpyHostToDevice, s));
  if (!on) HOR_HIP_TRY(hipMemcpyAsync(P->in_d0.p, d0, per * sizeof(double) * P->nt, hipMemcpyHostToDevice, s));
  HOR_TRY(hor_plan_eval_dev(P, P->in_dens.p, on ? nullptr : P->in_d0.p, P->out_buf.p));
  const size_t nout = (P->kind == HOR_P_OFF ? 1 : per) * (size_t)P->nt;
  HOR_HIP_TRY(hipMemcpyAsync(out, P->out_buf.p, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  if (!all_finite(out, nout)) return set_error(HOR_E_NAN, "non-finite result");
  return HOR_OK;
}

extern "C" int hor_plan_stats(hor_plan* P, double* direct_pairs, double* near_pairs, double* proxy_pairs,
                              double* moment_terms) {
  if (!P) return set_error(HOR_E_ARG, "null plan");
  if (direct_pairs) *direct_pairs = (double)P->ns * (double)P->nt;
  if (near_pairs) *near_pairs = P->method == HOR_DIRECT ? P->direct_pairs : P->near_pairs;
  if (proxy_pairs) *proxy_pairs = P->proxy_pairs;
  if (moment_terms) *moment_terms = P->moment_terms;
  return HOR_OK;
}

extern "C" int hor_plan_set_far_kernel(hor_plan* P, int kernel) { return hor::tree_plan_set_far_kernel(P, kernel); }

extern "C" int hor_plan_phase_timing(hor_plan* P, double* moments_ms, double* near_ms, double* far_ms) {
  if (!P) return set_error(HOR_E_ARG, "null plan");
  double ms[3];
  HOR_HIP_TRY(hipSetDevice(P->ctx->device));
  HOR_HIP_TRY(hipStreamSynchronize(P->ctx->stream));
  HOR_TRY(hor::tree_plan_phase_ms(P, ms));
  if (moments_ms) *moments_ms = ms[0];
  if (near_ms) *near_ms = ms[1];
  if (far_ms) *far_ms = ms[2];
  return HOR_OK;
}

// ---------------------------------------------------------------------------
// One-shot drop-ins

static int one_shot(hor_ctx* ctx, int kind, int method, const double* ctr, const hor_tree_params* tree,
                    double DEL, int ns, const double* sx, const double* sn, const double* sa,
                    const double* dens, int nt, const double* tx, const double* tb, const double* tx0,
                    const double* tn0, const double* d0, double* out) {
  hor_plan* P = nullptr;
  HOR_TRY(hor_plan_create(ctx, kind, method, HOR_FP64, DEL, ns, sx, sn, sa, nt, tx, tb, tx0, tn0, ctr,
                          tree, &P));
  int r = hor_plan_eval(P, dens, d0, out);
  hor_plan_destroy(P);
  return r;
}

extern "C" int hor_sl_offsurf(hor_ctx* ctx, double DEL, int ns, const double* sx, const double* sn,
                              const double* sa, const double* sf, int nt, const double* tx,
                              const double* tb, const double* tn0, const double* tf0, double* out) {
  return one_shot(ctx, HOR_SL_OFF, HOR_DIRECT, nullptr, nullptr, DEL, ns, sx, sn, sa, sf, nt, tx, tb,
                  nullptr, tn0, tf0, out);
}
extern "C" int hor_sl_onsurf(hor_ctx* ctx, double DEL, int ns, const double* sx, const double* sn,
                             const double* sa, const double* sf, int nt, const double* tx, double* out) {
  return one_shot(ctx, HOR_SL_ON, HOR_DIRECT, nullptr, nullptr, DEL, ns, sx, sn, sa, sf, nt, tx, nullptr,
                  nullptr, nullptr, nullptr, out);
}
extern "C" int hor_dl_offsurf(hor_ctx* ctx, double DEL, int ns, const double* sx, const double* sn,
                              const double* sa, const double* sg, int nt, const double* tx,
                              const double* tb, const double* tx0, const double* tn0, const double* tg0,
                              double* out) {
  return one_shot(ctx, HOR_DL_OFF, HOR_DIRECT, nullptr, nullptr, DEL, ns, sx, sn, sa, sg, nt, tx, tb, tx0,
                  tn0, tg0, out);
}
extern "C" int hor_dl_onsurf(hor_ctx* ctx, double DEL, int ns, const double* sx, const double* sn,
                             const double* sa, const double* sg, int nt, const double* tx, double* out) {
  return one_shot(ctx, HOR_DL_ON, HOR_DIRECT, nullptr, nullptr, DEL, ns, sx, sn, sa, sg, nt, tx, nullptr,
                  nullptr, nullptr, nullptr, out);
}
extern "C" int hor_sl_onsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree, double DEL,
                                  int ns, const double* sx, const double* sn, const double* sa,
                                  const double* sf, int nt, const double* tx, double* out) {
  return one_shot(ctx, HOR_SL_ON, HOR_TREE, ctr, tree, DEL, ns, sx, sn, sa, sf, nt, tx, nullptr, nullptr,
                  nullptr, nullptr, out);
}
extern "C" int hor_sl_offsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree, double DEL,
                                   int ns, const double* sx, const double* sn, const double* sa,
                                   const double* sf, int nt, const double* tx, const double* tb,
                                   const double* tn0, const double* tf0, double* out) {
  return one_shot(ctx, HOR_SL_OFF, HOR_TREE, ctr, tree, DEL, ns, sx, sn, sa, sf, nt, tx, tb, nullptr, tn0,
                  tf0, out);
}
extern "C" int hor_dl_onsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree, double DEL,
                                  int ns, const double* sx, const double* sn, const double* sa,
                                  const double* sg, int nt, const double* tx, double* out) {
  return one_shot(ctx, HOR_DL_ON, HOR_TREE, ctr, tree, DEL, ns, sx, sn, sa, sg, nt, tx, nullptr, nullptr,
                  nullptr, nullptr, out);
}
extern "C" int hor_dl_offsurf_tree(hor_ctx* ctx, const double* ctr, const hor_tree_params* tree, double DEL,
                                   int ns, const double* sx, const double* sn, const double* sa,
                                   const double* sg, int nt, const double* tx, const double* tb,
                                   const double* tx0, const double* tn0, const double* tg0, double* out) {
  return one_shot(ctx, HOR_DL_OFF, HOR_TREE, ctr, tree, DEL, ns, sx, sn, sa, sg, nt, tx, tb, tx0, tn0, tg0,
                  out);
}

// stokes_pressure_7ord (extend_to_grid/stokes.cpp:134-203) at every target.
extern "C" int hor_pressure_7ord(hor_ctx* ctx, double DEL, int ns, const double* sx, const double* sn,
                                 const double* sa, const double* sf, int nt, const double* tx, const double* tb,
                                 const double* tn0, const double* tf0, double* out) {
  return one_shot(ctx, HOR_P_OFF, HOR_DIRECT, nullptr, nullptr, DEL, ns, sx, sn, sa, sf, nt, tx, tb, nullptr,
                  tn0, tf0, out);
}

// stokes_pressure_away (extend_to_grid/stokes.cpp:102-127): the pressure
// operator with every target beyond 8 delta (plain 1/r^3 for every pair) and
// no subtraction (f0 = 0), which is 1/(4 pi) sum A f.dx / r^3.
extern "C" int hor_pressure_away(hor_ctx* ctx, int ns, const double* sx, const double* sn, const double* sa,
                                 const double* sf, int nt, const double* tx, double* out) {
  if (nt <= 0) return set_error(HOR_E_ARG, "hor_pressure_away: need nt > 0");
  std::vector<double> far_b((size_t)nt, 1e300), zero(3 * (size_t)nt, 0.0);
  return one_shot(ctx, HOR_P_OFF, HOR_DIRECT, nullptr, nullptr, 1.0, ns, sx, sn, sa, sf, nt, tx, far_b.data(),
                  nullptr, zero.data(), zero.data(), out);
}
