// hor_dl_onsurf_dev_noalloc: the on-surface stresslet sum of hor_dl_onsurf
// (eval_Stokes_DL_onSurf, highreg_stokes_2surf/stresslets.cpp:186-228) on
// device pointers, with no device allocation.  Every intermediate buffer the
// direct sum needs (SoA copies of the sources, source records, subtile and
// wave boxes, per-target data, the chunk partial sums and the kernel table)
// is carved out of one file-scope device array, `scratch`, of kScratchBytes.
//
// Differences from the plan path: sources and targets stay in the caller's
// order (no host sort), so the 64-source subtile boxes are looser and fewer
// subtiles take the plain-kernel fast path; the pair arithmetic, the chunk
// partial sums and the finalize are the plan path's own kernels, and the
// result differs from hor_dl_onsurf's only by the order of the sums.
//
// The array lives in the library's code object: each GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// at the entry point's first call on a device, with that device current,
// hipGetSymbolAddress gives that device's address, which is kept per device
// and used for every launch argument and copy.  Every call on a GPU shares
// that GPU's array and leaves its contents behind; nothing in it must start at
// zero (the sum kernel overwrites every partial sum it later reads), and calls
// that overlap on one GPU (two contexts, two threads) race on it.
#include <cmath>
#include <cstring>
#include <mutex>

#include "device.h"
#include "direct.h"
#include "plan.h"

namespace hor {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per device
constexpr int kNoAllocMaxChunks = 8;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};   // per device: the address of `scratch` on it

// The device address of `scratch` on the current device (resolved once per
// device), with the on-surface stresslet table copied to its first bytes.
int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(HOR_E_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    HOR_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(HOR_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    HOR_HIP_TRY(hipMemcpy(p, kDlOnTable, sizeof(double) * kDlOnTabLen, hipMemcpyHostToDevice));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return HOR_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the regions of one call (sizes from ns, nt).
struct Layout {
  int nsp, nsub, ntp, nchunk, subs_per_chunk;
  size_t tab, sx, sn, sa, sd, rec, sub_box, sperm, tx, tcoef, tnear, wave_box, xi, tperm, tsub, partial, total;
};

Layout layout(int ns, int nt) {
  Layout L;
  L.nsp = std::max(kWave, (ns + kWave - 1) / kWave * kWave);
  L.nsub = (ns + kWave - 1) / kWave;
  L.ntp = std::max(kBlock, (nt + kBlock - 1) / kBlock * kBlock);
  // source chunks as the plan path chooses them, at most kNoAllocMaxChunks
  const int nblk = L.ntp / kBlock;
  const int want = std::max(1, (40 * 2048 + nblk - 1) / nblk);
  const int nchunk = std::min({want, std::max(1, L.nsub / 8), kNoAllocMaxChunks});
  L.subs_per_chunk = (L.nsub + nchunk - 1) / nchunk;
  L.nchunk = (L.nsub + L.subs_per_chunk - 1) / L.subs_per_chunk;
  const size_t d = sizeof(double);
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = align256(o + bytes); return at; };
  L.tab = take(d * kDlOnTabLen);   // first: copied once at set-up, never overwritten
  L.sx = take(d * 3 * L.nsp);
  L.sn = take(d * 3 * L.nsp);
  L.sa = take(d * L.nsp);
  L.sd = take(d * 3 * L.nsp);
  L.rec = take(sizeof(SrcRec) * L.nsp);
  L.sub_box = take(d * 6 * std::max(L.nsub, 1));
  L.sperm = take(sizeof(int) * L.nsp);
  L.tx = take(d * 3 * L.ntp);
  L.tcoef = take(d * 3 * L.ntp);
  L.tnear = take(sizeof(int) * L.ntp);
  L.wave_box = take(d * 8 * (L.ntp / kWave));
  L.xi = take(d * L.ntp);
  L.tperm = take(sizeof(int) * L.ntp);
  L.tsub = take(d * 3 * L.ntp);
  L.partial = take(d * 3 * (size_t)L.ntp * L.nchunk);
  L.total = o;
  return L;
}

struct Coefs {
  double c[3];   // regularization coefficients c1, c2, c3 (target_coefs)
};

// Sources in the caller's order as SoA with stride nsp (zero padding).
__global__ void k_na_sources(int ns, int nsp, const double* sx, const double* sn, const double* sa, double* X,
                             double* N, double* A, int* perm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsp) return;
  const bool in = j < ns;
  for (int k = 0; k < 3; k++) {
    X[(size_t)k * nsp + j] = in ? sx[3 * (size_t)j + k] : 0.0;
    N[(size_t)k * nsp + j] = in ? sn[3 * (size_t)j + k] : 0.0;
  }
  A[j] = in ? sa[j] : 0.0;
  perm[j] = j;
}

// Targets in the caller's order, padded to ntp with copies of the last one:
// positions (SoA), zero regularization coefficients (the on-surface stresslet
// takes its kernel from its table), class "near", jump 1/2.
__global__ void k_na_targets(int nt, int ntp, const double* tx, Coefs c, double* X, double* C, int* nearp,
                             double* xi, int* tperm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp) return;
  const int o = min(i, nt - 1);
  for (int k = 0; k < 3; k++) X[(size_t)k * ntp + i] = tx[3 * (size_t)o + k];
  for (int k = 0; k < 3; k++) C[(size_t)k * ntp + i] = c.c[k];
  nearp[i] = 1;
  xi[i] = i < nt ? 0.5 : 0.0;   // stresslets.cpp:276-278
  tperm[i] = o;
}

// Bounding box of each group of 64 points of a SoA array (stride np, the
// first n valid): 6 values per group (lo xyz, hi xyz), or 8 with a zero
// "all far" flag and padding (wave boxes of on-surface targets).
__global__ void k_na_boxes(int n, int np, int ngroups, const double* X, double* box, int stride) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  const int j1 = min(n, (g + 1) * kWave);
  for (int j = g * kWave; j < j1; j++)
    for (int k = 0; k < 3; k++) {
      const double v = X[(size_t)k * np + j];
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  for (int k = 0; k < 3; k++) {
    box[(size_t)stride * g + k] = lo[k];
    box[(size_t)stride * g + 3 + k] = hi[k];
  }
  if (stride == 8) {
    box[(size_t)stride * g + 6] = 0.0;
    box[(size_t)stride * g + 7] = 0.0;
  }
}

}  // namespace
}  // namespace hor

using namespace hor;

extern "C" int hor_dl_onsurf_dev_noalloc(hor_ctx* ctx, double DEL, int ns, const double* sx_dev,
                                         const double* sn_dev, const double* sa_dev, const double* sg_dev, int nt,
                                         const double* tx_dev, double* out_dev) {
  if (!ctx || !sx_dev || !sn_dev || !sa_dev || !sg_dev || !tx_dev || !out_dev)
    return set_error(HOR_E_ARG, "hor_dl_onsurf_dev_noalloc: null argument");
  if (!(DEL > 0) || !std::isfinite(DEL)) return set_error(HOR_E_ARG, "hor_dl_onsurf_dev_noalloc: DEL must be finite and > 0");
  if (ns <= 0 || nt <= 0 || nt > ns)
    return set_error(HOR_E_ARG, "hor_dl_onsurf_dev_noalloc: need 0 < nt <= ns (target i is source i), got ns=%d nt=%d",
                     ns, nt);
  const Layout L = layout(ns, nt);
  if (L.total > kScratchBytes)
    return set_error(HOR_E_ARG,
                     "hor_dl_onsurf_dev_noalloc: ns=%d, nt=%d need %zu bytes of scratch, more than its %zu "
                     "(ns = nt = 62000 fit); use hor_dl_onsurf or a plan",
                     ns, nt, L.total, kScratchBytes);
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  double* base = nullptr;
  HOR_TRY(scratch_on(ctx->device, &base));
  char* b = reinterpret_cast<char*>(base);
  auto D = [&](size_t off) { return reinterpret_cast<double*>(b + off); };
  auto I = [&](size_t off) { return reinterpret_cast<int*>(b + off); };
  hipStream_t s = ctx->stream;
  const dim3 blk(256);
  hipLaunchKernelGGL(k_na_sources, dim3((L.nsp + 255) / 256), blk, 0, s, ns, L.nsp, sx_dev, sn_dev, sa_dev, D(L.sx),
                     D(L.sn), D(L.sa), I(L.sperm));
  hipLaunchKernelGGL(k_na_boxes, dim3((L.nsub + 255) / 256), blk, 0, s, ns, L.nsp, L.nsub, D(L.sx), D(L.sub_box), 6);
  Coefs c;
  target_coefs(HOR_DL_ON, true, 0.0, DEL, c.c);
  hipLaunchKernelGGL(k_na_targets, dim3((L.ntp + 255) / 256), blk, 0, s, nt, L.ntp, tx_dev, c, D(L.tx), D(L.tcoef),
                     I(L.tnear), D(L.xi), I(L.tperm));
  const int nw = L.ntp / kWave;
  hipLaunchKernelGGL(k_na_boxes, dim3((nw + 255) / 256), blk, 0, s, L.ntp, L.ntp, nw, D(L.tx), D(L.wave_box), 8);
  HOR_HIP_TRY(hipGetLastError());
  // the plan path's kernels on the scratch regions
  HOR_TRY(launch_gather3(sg_dev, I(L.sperm), ns, L.nsp, D(L.sx), D(L.sn), D(L.sa), D(L.sd),
                         reinterpret_cast<SrcRec*>(b + L.rec), nullptr, D(L.sub_box), 0, s));
  HOR_TRY(launch_target_sub(3, nt, L.ntp, I(L.tperm), I(L.tnear), nullptr, nullptr, sg_dev, nullptr, D(L.tsub), s));
  DirectArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ns = ns; A.nsp = L.nsp; A.nt = nt; A.ntp = L.ntp;
  A.nsub = L.nsub; A.subs_per_chunk = L.subs_per_chunk;
  A.DEL = DEL;
  const double fd = 8.0 * DEL * (1.0 + 1e-10);
  A.far_d2 = fd * fd;
  A.near_r2 = 64.0 * DEL * DEL * (1.0 + 1e-12);
  A.near64_r2 = 16.0 * DEL * DEL;
  A.sx = D(L.sx); A.sn = D(L.sn); A.sa = D(L.sa); A.sd = D(L.sd); A.sub_box = D(L.sub_box);
  A.rec = reinterpret_cast<const SrcRec*>(b + L.rec);
  A.tx = D(L.tx); A.tcoef = D(L.tcoef); A.tsub = D(L.tsub);
  A.tnear = I(L.tnear); A.wave_box = D(L.wave_box); A.partial = D(L.partial);
  A.dlon_tab = D(L.tab);
  HOR_HIP_TRY(hipEventRecord(ctx->ev[2], s));
  HOR_TRY(launch_direct(A, HOR_DL_ON, false, L.nchunk, s));
  HOR_HIP_TRY(hipEventRecord(ctx->ev[3], s));
  HOR_TRY(launch_finalize(D(L.partial), L.nchunk, nt, L.ntp, op_scale(HOR_DL_ON), D(L.xi), D(L.tsub), I(L.tperm),
                          out_dev, nullptr, 1.0, 0, s));
  HOR_HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  HOR_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
  ctx->last_kernel_ms = ms;
  return HOR_OK;
}
