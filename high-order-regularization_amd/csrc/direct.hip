// Direct regularized Stokeslet / stresslet sums on gfx950.
//
// One thread owns one target (targets sorted far-first, then by Morton key,
// so a wave's 64 targets are spatially compact and share one regularization
// class).  Every lane of a wave visits the same source at the same time, so
// the source records (Morton sorted, array of structures) stream through
// wave-uniform scalar loads into SGPRs and feed the FP64 VALU directly: no LDS
// traffic, no barriers.  A wave walks the sources in 64-source subtiles and,
// per subtile, takes one of two wave-uniform paths:
//   * fast path: the subtile's bounding box lies beyond 8*delta (plus a
//     rounding margin) of the wave's target box, or every target of the wave
//     is beyond 8*delta of the surface: every pair uses the plain kernel
//     (utilities.cpp:186-190, 265-269), no erf/exp, no branch;
//   * mixed path: per pair, self limit / plain / regularized exactly as
//     Stokes_SL_7ord / Stokes_DL_7ord (utilities.cpp:179-323).
// Sources are split into chunks across grid.y when there are too few target
// workgroups to fill the 256 CUs; each chunk writes a partial sum and a
// finalize kernel adds the chunks in a fixed order (deterministic).
#include "pairs.h"

namespace hor {

// LDS bytes of a sum kernel: its table (stage_table).
inline size_t lds_bytes(int kind) { return sizeof(double) * table_len(kind); }

template <int KIND, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_direct(DirectArgs A) {
  const double* tab = stage_table<KIND>(A);
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kBlock + tid;     // sorted target (ntp is a multiple of kBlock)
  const int wave = i / kWave;
  Tgt t;
  load_target<KIND>(A, i, t);
  const double* wb = A.wave_box + 8 * wave;
  const bool all_far = wb[6] != 0.0;
  double u[3] = {0.0, 0.0, 0.0};
  // this chunk's 64-source subtiles
  const int sub0 = blockIdx.y * A.subs_per_chunk;
  const int sub1 = min(A.nsub, sub0 + A.subs_per_chunk);
  range_sum<KIND, MIXED, true>(A, sub0 * kWave, min(A.ns, sub1 * kWave), sub0, wb, all_far, t, tab, u);
  double* P = A.partial + (size_t)blockIdx.y * 3 * A.ntp;
  P[i] = u[0];
  P[A.ntp + i] = u[1];
  P[2 * A.ntp + i] = u[2];
}

// The regularized pressure (HOR_P_OFF, FP64), one value per target and
// chunk.  Launch bounds for five waves per SIMD (at most 96 VGPRs): the
// per-subtile target data of range_sum_p would otherwise take 103 VGPRs and
// four waves, 3% slower on the band workload.
__global__ __launch_bounds__(kBlock, 5) void k_pressure(DirectArgs A) {
  const double* tab = stage_table<HOR_P_OFF>(A);
  const int i = blockIdx.x * kBlock + threadIdx.x;   // sorted target (ntp is a multiple of kBlock)
  Tgt t;
  load_target<HOR_P_OFF>(A, i, t);
  const double* wb = A.wave_box + 8 * (i / kWave);
  const bool all_far = wb[6] != 0.0;
  double u[3] = {0.0, 0.0, 0.0};
  const int sub0 = blockIdx.y * A.subs_per_chunk;
  const int sub1 = min(A.nsub, sub0 + A.subs_per_chunk);
  range_sum_p(A, sub0 * kWave, min(A.ns, sub1 * kWave), sub0, wb, all_far, t, tab, u);
  A.partial[(size_t)blockIdx.y * 3 * A.ntp + i] = u[0];
}

// Stokeslet + stresslet at off-surface targets in one pass (HOR_SLDL_OFF):
// the partial sums of both kernels, 6 per target per chunk.
template <bool MIXED>
__global__ __launch_bounds__(kBlock, 4) void k_direct_sldl(DirectArgs A) {
  const double* tab = stage_table<HOR_SL_OFF>(A);
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kBlock + tid;     // sorted target (ntp is a multiple of kBlock)
  const int wave = i / kWave;
  Tgt t;
  load_target<HOR_SLDL_OFF>(A, i, t);
  const double* wb = A.wave_box + 8 * wave;
  const bool all_far = wb[6] != 0.0;
  double us[3] = {0.0, 0.0, 0.0}, ud[3] = {0.0, 0.0, 0.0};
  const int sub0 = blockIdx.y * A.subs_per_chunk;
  const int sub1 = min(A.nsub, sub0 + A.subs_per_chunk);
  range_sum_sldl<MIXED>(A, sub0 * kWave, min(A.ns, sub1 * kWave), sub0, wb, all_far, t, tab, us, ud);
  double* P = A.partial + (size_t)blockIdx.y * 6 * A.ntp;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    P[(size_t)k * A.ntp + i] = us[k];
    P[(size_t)(3 + k) * A.ntp + i] = ud[k];
  }
}

// The two results of the fused sum in the caller's order: out[0, 3nt) =
// scale_sl*sum (the SL_OFF finalize), out[3nt, 6nt) = scale_dl*sum + xi*g0
// (the DL_OFF finalize), chunk partials added in chunk order.
__global__ void k_sldl_finalize(const double* partial, int nchunk, int nt, int ntp, double scale_sl,
                                double scale_dl, const double* xi, const double* g0, const int* tperm,
                                double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  int o = tperm[i];
  for (int k = 0; k < 3; k++) {
    double s = 0.0, d = 0.0;
    for (int c = 0; c < nchunk; c++) {
      s += partial[((size_t)c * 6 + k) * ntp + i];
      d += partial[((size_t)c * 6 + 3 + k) * ntp + i];
    }
    out[3 * o + k] = s * scale_sl;
    out[3 * (size_t)nt + 3 * o + k] = d * scale_dl + xi[i] * g0[(size_t)k * ntp + i];
  }
}

// Pressure: one value per target, v = scale*sum + xi*(f0.n0) with xi = -1
// inside (b < -1e-12, stokes.cpp:198-200), combined into out like
// k_direct_finalize.
__global__ void k_pressure_finalize(const double* partial, int nchunk, int nt, int ntp, double scale,
                                    const double* xi, const double* f0n0, const int* tperm, double* out,
                                    const double* base, double coef, int accumulate) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  int o = tperm[i];
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += partial[(size_t)c * 3 * ntp + i];
  double w = coef * (s * scale + xi[i] * f0n0[i]);
  if (accumulate) out[o] += w;
  else if (base) out[o] = base[o] + w;
  else out[o] = w;
}

// Adds the chunk partials in chunk order, scales, adds the jump term and
// writes v = scale*sum + xi*g0 in the caller's order (xyz interleaved):
//   out = coef*v                 (accumulate = 0, base = NULL)
//   out = base + coef*v          (accumulate = 0; the solver's u - 2/3 DL)
//   out = out + coef*v           (accumulate = 1)
__global__ void k_direct_finalize(const double* partial, int nchunk, int nt, int ntp, double scale,
                                  const double* xi, const double* g0, const int* tperm, double* out,
                                  const double* base, double coef, int accumulate) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  int o = tperm[i];
  for (int k = 0; k < 3; k++) {
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += partial[((size_t)c * 3 + k) * ntp + i];
    double v = s * scale;
    if (xi) v += xi[i] * g0[(size_t)k * ntp + i];
    double w = coef * v;
    if (accumulate) out[3 * o + k] += w;
    else if (base) out[3 * o + k] = base[3 * o + k] + w;
    else out[3 * o + k] = w;
  }
}

// Per-evaluation source records: sorted AoS (position, weight, normal,
// density) for the sum kernels, plus the density as sorted SoA for the
// treecode moments: rec[j] <- geometry[j], dens[3*perm[j] + k].
__global__ void k_gather3(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                          const double* sa, double* sd, SrcRec* rec, SrcRec32* rec32, const double* sub_box,
                          int scale_density, const int* sub_of) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int o = perm[j];
  SrcRec r;
  r.x = sx[j]; r.y = sx[np + j]; r.z = sx[2 * np + j]; r.a = sa[j];
  r.nx = r.a * sn[j]; r.ny = r.a * sn[np + j]; r.nz = r.a * sn[2 * np + j]; r.pad0 = 0.0;
  r.d0 = dens[3 * o]; r.d1 = dens[3 * o + 1]; r.d2 = dens[3 * o + 2]; r.pad1 = 0.0;
  sd[j] = r.d0;
  sd[np + j] = r.d1;
  sd[2 * np + j] = r.d2;
  if (scale_density) {
    r.d0 *= r.a;
    r.d1 *= r.a;
    r.d2 *= r.a;
  }
  rec[j] = r;
  if (rec32) {
    // positions relative to the centre of the source's subtile box (the FP32
    // plain loops: the direct sums' 64-source subtiles, the treecode's leaf
    // subtiles through sub_of), absolute otherwise
    double c[3] = {0.0, 0.0, 0.0};
    if (sub_box) {
      const double* b = sub_box + 6 * (sub_of ? sub_of[j] : j / kWave);
      for (int k = 0; k < 3; k++) c[k] = 0.5 * (b[k] + b[3 + k]);
    }
    SrcRec32 q;
    q.x = (float)(r.x - c[0]); q.y = (float)(r.y - c[1]); q.z = (float)(r.z - c[2]); q.a = (float)r.a;
    q.nx = (float)r.nx; q.ny = (float)r.ny; q.nz = (float)r.nz; q.pad0 = 0.0f;
    q.d0 = (float)r.d0; q.d1 = (float)r.d1; q.d2 = (float)r.d2; q.pad1 = 0.0f;
    rec32[j] = q;
  }
}

// Pressure records: recp[j] <- (y', a n.y', a n, a f.y', a f, a n x y', |y'|^2)
// of sorted source j with y' = y - c, c the centre of its 64-subtile box,
// dens in the caller's order.
__global__ void k_gather_p(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                           const double* sa, const double* sub_box, SrcRecP* recp) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int o = perm[j];
  const double a = sa[j];
  const double* bx = sub_box + 6 * (j / kWave);
  SrcRecP r;
  r.x = sx[j] - 0.5 * (bx[0] + bx[3]);
  r.y = sx[np + j] - 0.5 * (bx[1] + bx[4]);
  r.z = sx[2 * np + j] - 0.5 * (bx[2] + bx[5]);
  r.nx = a * sn[j]; r.ny = a * sn[np + j]; r.nz = a * sn[2 * np + j];
  r.d0 = a * dens[3 * o]; r.d1 = a * dens[3 * o + 1]; r.d2 = a * dens[3 * o + 2];
  r.al = r.nx * r.x + r.ny * r.y + r.nz * r.z;
  r.be = r.d0 * r.x + r.d1 * r.y + r.d2 * r.z;
  r.m0 = r.ny * r.z - r.nz * r.y;
  r.m1 = r.nz * r.x - r.nx * r.z;
  r.m2 = r.nx * r.y - r.ny * r.x;
  r.xx = r.x * r.x + r.y * r.y + r.z * r.z;
  r.pad = 0.0;
  recp[j] = r;
}

// Fused records (HOR_SLDL_OFF), dens = [f | g]: recv[j] <- (y, a n, a f, g)
// of sorted source j, the SL and DL records' values; with recv32 also their
// FP32 copies, y relative to the centre of the source's 64-subtile box.
__global__ void k_gather_sldl(const double* dens, const int* perm, int n, int np, const double* sx,
                              const double* sn, const double* sa, SrcRecV* recv, SrcRecV32* recv32,
                              const double* sub_box) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int o = perm[j];
  const double a = sa[j];
  const double* g = dens + 3 * (size_t)n;
  SrcRecV r;
  r.x = sx[j]; r.y = sx[np + j]; r.z = sx[2 * np + j];
  r.nx = a * sn[j]; r.ny = a * sn[np + j]; r.nz = a * sn[2 * np + j];
  r.f0 = dens[3 * o] * a; r.f1 = dens[3 * o + 1] * a; r.f2 = dens[3 * o + 2] * a;
  r.g0 = g[3 * o]; r.g1 = g[3 * o + 1]; r.g2 = g[3 * o + 2];
  recv[j] = r;
  if (recv32) {
    const double* bx = sub_box + 6 * (j / kWave);
    SrcRecV32 q;
    q.x = (float)(r.x - 0.5 * (bx[0] + bx[3]));
    q.y = (float)(r.y - 0.5 * (bx[1] + bx[4]));
    q.z = (float)(r.z - 0.5 * (bx[2] + bx[5]));
    q.nx = (float)r.nx; q.ny = (float)r.ny; q.nz = (float)r.nz;
    q.f0 = (float)r.f0; q.f1 = (float)r.f1; q.f2 = (float)r.f2;
    q.g0 = (float)r.g0; q.g1 = (float)r.g1; q.g2 = (float)r.g2;
    recv32[j] = q;
  }
}

// Per-target subtraction data (sorted order):
//   mode 0 (SL_OFF): sub = near ? d0.n0 : 0       (stokeslets.cpp:72)
//   mode 1 (SL_ON):  sub = f_o . n_o of source o  (stokeslets.cpp:167)
//   mode 2 (DL_OFF): g0 = near ? d0 : 0            (leaf / direct sums)
//   mode 3 (DL_ON):  g0 = g_o of source o          (stresslets.cpp:212-214)
//   mode 4 (DL_OFF): g0 = d0 for every target      (treecode proxies and jump)
//   mode 5 (P_OFF):  [d0.n0, n0 x d0] for every target (stokes.cpp:146-147)
//   mode 6 (SLDL_OFF, d0 = [f0 | g0]): [near ? f0.n0 : 0, near ? g0 : 0]
__global__ void k_target_sub(int mode, int nt, int ntp, const int* tperm, const int* tnear,
                             const double* d0, const double* n0_orig, const double* dens,
                             const double* sn_orig, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp) return;
  if (i >= nt) {
    out[i] = 0;
    if (mode >= 2) { out[ntp + i] = 0; out[2 * ntp + i] = 0; }
    if (mode == 5 || mode == 6) out[3 * ntp + i] = 0;
    return;
  }
  int o = tperm[i];
  bool nearp = tnear[i] != 0;
  if (mode == 0) {
    out[i] = nearp ? d0[3 * o] * n0_orig[3 * o] + d0[3 * o + 1] * n0_orig[3 * o + 1] +
                         d0[3 * o + 2] * n0_orig[3 * o + 2]
                   : 0.0;
  } else if (mode == 1) {
    out[i] = dens[3 * o] * sn_orig[3 * o] + dens[3 * o + 1] * sn_orig[3 * o + 1] +
             dens[3 * o + 2] * sn_orig[3 * o + 2];
  } else if (mode == 2) {
    for (int k = 0; k < 3; k++) out[(size_t)k * ntp + i] = nearp ? d0[3 * o + k] : 0.0;
  } else if (mode == 3) {
    for (int k = 0; k < 3; k++) out[(size_t)k * ntp + i] = dens[3 * o + k];
  } else if (mode == 6) {
    out[i] = nearp ? d0[3 * o] * n0_orig[3 * o] + d0[3 * o + 1] * n0_orig[3 * o + 1] +
                         d0[3 * o + 2] * n0_orig[3 * o + 2]
                   : 0.0;
    for (int k = 0; k < 3; k++) out[(size_t)(1 + k) * ntp + i] = nearp ? d0[3 * (size_t)nt + 3 * o + k] : 0.0;
  } else if (mode == 5) {
    const double* f = d0 + 3 * o;
    const double* n = n0_orig + 3 * o;
    out[i] = f[0] * n[0] + f[1] * n[1] + f[2] * n[2];
    out[ntp + i] = n[1] * f[2] - n[2] * f[1];
    out[2 * (size_t)ntp + i] = n[2] * f[0] - n[0] * f[2];
    out[3 * (size_t)ntp + i] = n[0] * f[1] - n[1] * f[0];
  } else {
    for (int k = 0; k < 3; k++) out[(size_t)k * ntp + i] = d0[3 * o + k];
  }
}

// SL_OFF proxy subtraction uses f0.n0 for every target (KITC.cpp:1568).
__global__ void k_target_f0n0_all(int nt, int ntp, const int* tperm, const double* d0,
                                  const double* n0_orig, double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp) return;
  if (i >= nt) { out[i] = 0; return; }
  int o = tperm[i];
  out[i] = d0[3 * o] * n0_orig[3 * o] + d0[3 * o + 1] * n0_orig[3 * o + 1] + d0[3 * o + 2] * n0_orig[3 * o + 2];
}

// ---------------------------------------------------------------------------
// Host launchers.

int launch_direct(const DirectArgs& A, int kind, bool mixed, int nchunk, hipStream_t s) {
  dim3 grid(A.ntp / kBlock, nchunk), block(kBlock);
  if (kind == HOR_SLDL_OFF) {
    if (mixed) hipLaunchKernelGGL(k_direct_sldl<true>, grid, block, lds_bytes(HOR_SL_OFF), s, A);
    else hipLaunchKernelGGL(k_direct_sldl<false>, grid, block, lds_bytes(HOR_SL_OFF), s, A);
    HOR_HIP_TRY(hipGetLastError());
    return HOR_OK;
  }
#define HOR_LAUNCH(K, M) hipLaunchKernelGGL((k_direct<K, M>), grid, block, lds_bytes(K), s, A)
  if (!mixed) {
    switch (kind) {
      case HOR_SL_ON: HOR_LAUNCH(HOR_SL_ON, false); break;
      case HOR_SL_OFF: HOR_LAUNCH(HOR_SL_OFF, false); break;
      case HOR_DL_ON: HOR_LAUNCH(HOR_DL_ON, false); break;
      case HOR_P_OFF: hipLaunchKernelGGL(k_pressure, grid, block, lds_bytes(HOR_P_OFF), s, A); break;
      default: HOR_LAUNCH(HOR_DL_OFF, false); break;
    }
  } else {
    if (kind == HOR_P_OFF) return set_error(HOR_E_ARG, "the pressure operator is FP64 only");
    switch (kind) {
      case HOR_SL_ON: HOR_LAUNCH(HOR_SL_ON, true); break;
      case HOR_SL_OFF: HOR_LAUNCH(HOR_SL_OFF, true); break;
      case HOR_DL_ON: HOR_LAUNCH(HOR_DL_ON, true); break;
      default: HOR_LAUNCH(HOR_DL_OFF, true); break;
    }
  }
#undef HOR_LAUNCH
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_finalize(const double* partial, int nchunk, int nt, int ntp, double scale, const double* xi,
                    const double* g0, const int* tperm, double* out, const double* base,
                    double coef, int accumulate, hipStream_t s) {
  if (nt == 0) return HOR_OK;
  hipLaunchKernelGGL(k_direct_finalize, dim3((nt + 255) / 256), dim3(256), 0, s, partial, nchunk, nt,
                     ntp, scale, xi, g0, tperm, out, base, coef, accumulate);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_sldl_finalize(const double* partial, int nchunk, int nt, int ntp, double scale_sl, double scale_dl,
                         const double* xi, const double* g0, const int* tperm, double* out, hipStream_t s) {
  if (nt == 0) return HOR_OK;
  hipLaunchKernelGGL(k_sldl_finalize, dim3((nt + 255) / 256), dim3(256), 0, s, partial, nchunk, nt, ntp, scale_sl,
                     scale_dl, xi, g0, tperm, out);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_gather_sldl(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                       const double* sa, SrcRecV* recv, SrcRecV32* recv32, const double* sub_box, hipStream_t s) {
  if (n == 0) return HOR_OK;
  hipLaunchKernelGGL(k_gather_sldl, dim3((n + 255) / 256), dim3(256), 0, s, dens, perm, n, np, sx, sn, sa, recv,
                     recv32, sub_box);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_gather_p(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                    const double* sa, const double* sub_box, SrcRecP* recp, hipStream_t s) {
  if (n == 0) return HOR_OK;
  hipLaunchKernelGGL(k_gather_p, dim3((n + 255) / 256), dim3(256), 0, s, dens, perm, n, np, sx, sn, sa, sub_box,
                     recp);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_pressure_finalize(const double* partial, int nchunk, int nt, int ntp, double scale, const double* xi,
                             const double* f0n0, const int* tperm, double* out, const double* base, double coef,
                             int accumulate, hipStream_t s) {
  if (nt == 0) return HOR_OK;
  hipLaunchKernelGGL(k_pressure_finalize, dim3((nt + 255) / 256), dim3(256), 0, s, partial, nchunk, nt, ntp, scale,
                     xi, f0n0, tperm, out, base, coef, accumulate);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_gather3(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                   const double* sa, double* sd, SrcRec* rec, SrcRec32* rec32, const double* sub_box,
                   int scale_density, hipStream_t s, const int* sub_of) {
  if (n == 0) return HOR_OK;
  hipLaunchKernelGGL(k_gather3, dim3((n + 255) / 256), dim3(256), 0, s, dens, perm, n, np, sx, sn, sa, sd, rec, rec32,
                     sub_box, scale_density, sub_of);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_target_sub(int mode, int nt, int ntp, const int* tperm, const int* tnear, const double* d0,
                      const double* n0_orig, const double* dens, const double* sn_orig, double* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_target_sub, dim3((ntp + 255) / 256), dim3(256), 0, s, mode, nt, ntp, tperm,
                     tnear, d0, n0_orig, dens, sn_orig, out);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int launch_target_f0n0_all(int nt, int ntp, const int* tperm, const double* d0, const double* n0_orig,
                           double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_target_f0n0_all, dim3((ntp + 255) / 256), dim3(256), 0, s, nt, ntp, tperm, d0,
                     n0_orig, out);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

}  // namespace hor
