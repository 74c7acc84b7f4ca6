// Kernel argument blocks and launchers of the sum kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "hor_internal.h"

namespace hor {

// Direct sums / treecode leaf sums.  All device pointers.  Sources are in the
// plan's sorted order as structure-of-arrays with stride nsp; targets in the
// plan's sorted order with stride ntp (a multiple of 256).
struct DirectArgs {
  int ns, nsp, nt, ntp;
  int nsub, subs_per_chunk;        // 64-source subtiles, subtiles per grid.y chunk
  double DEL, far_d2;              // delta, (8 delta (1+1e-10))^2
  double near_r2;                  // (8 delta)^2 (1+1e-12): pairs beyond are plain for sure
  double near64_r2;                // (4 delta)^2: mixed precision keeps pairs within it in FP64
  const double* sx;                // 3*nsp positions
  const double* sn;                // 3*nsp normals
  const double* sa;                // nsp weights
  const double* sd;                // 3*nsp density (f or g)
  const struct SrcRec* rec;        // nsp source records (sorted), this evaluation's density
  const struct SrcRecP* recp;      // P_OFF: nsp pressure records
  const struct SrcRecV* recv;      // SLDL_OFF: nsp fused records
  const struct SrcRecV32* recv32;  // SLDL_OFF mixed precision: FP32 copies (y - subtile centre)
  const struct SrcRec32* rec32;    // FP32 copies (mixed precision) or null; direct plans store
                                   // y - (centre of the source's 64-subtile box)
  const double* sub_box;           // 6 per 64-source subtile (lo xyz, hi xyz)
  const double* tx;                // 3*ntp positions
  const double* tcoef;             // 3*ntp regularization coefficients c1, c2, c3 (zero beyond 8 delta)
  const double* tsub;              // SL: ntp f0.n0; DL: 3*ntp g0; P: ntp f0.n0, 3*ntp n0 x f0;
                                   // SLDL: ntp f0.n0, 3*ntp g0
  const double* tx0;               // DL_OFF: 3*ntp nearest point
  const double* tn0;               // DL_OFF: 3*ntp nearest normal
  const double* tb;                // DL_OFF: ntp signed distance
  const int* tnear;                // ntp: 1 = regularized target (|b|/delta <= 8 or on surface)
  const double* wave_box;          // 8 per 64-target group: lo xyz, hi xyz, all-far flag, pad
  double* partial;                 // nchunk*3*ntp partial sums (nchunk*6*ntp for SLDL)
  const double* dlon_tab;          // kDlOnTable: the on-surface stresslet kernel (staged in LDS)
  const double* ae_tab;            // kErfExpTable: erf(rho)/rho and exp(-rho^2) (staged in LDS)
};

int launch_direct(const DirectArgs& A, int kind, bool mixed, int nchunk, hipStream_t s);
int launch_finalize(const double* partial, int nchunk, int nt, int ntp, double scale,
                    const double* xi, const double* g0, const int* tperm, double* out,
                    const double* base, double coef, int accumulate, hipStream_t s);
int launch_pressure_finalize(const double* partial, int nchunk, int nt, int ntp, double scale, const double* xi,
                             const double* f0n0, const int* tperm, double* out, const double* base, double coef,
                             int accumulate, hipStream_t s);
// rec32 positions are relative to the centre of box sub_of[j] of sub_box
// (sub_of NULL: box j/64), or absolute when sub_box is NULL.
int launch_gather3(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                   const double* sa, double* sd, struct SrcRec* rec, struct SrcRec32* rec32,
                   const double* sub_box, int scale_density, hipStream_t s, const int* sub_of = nullptr);
int launch_sldl_finalize(const double* partial, int nchunk, int nt, int ntp, double scale_sl, double scale_dl,
                         const double* xi, const double* g0, const int* tperm, double* out, hipStream_t s);
// Pressure records (SrcRecP), positions relative to their subtile's box centre.
int launch_gather_p(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                    const double* sa, const double* sub_box, struct SrcRecP* recp, hipStream_t s);
int launch_gather_sldl(const double* dens, const int* perm, int n, int np, const double* sx, const double* sn,
                       const double* sa, struct SrcRecV* recv, struct SrcRecV32* recv32, const double* sub_box,
                       hipStream_t s);
int launch_target_sub(int mode, int nt, int ntp, const int* tperm, const int* tnear,
                      const double* d0, const double* n0_orig, const double* dens,
                      const double* sn_orig, double* out, hipStream_t s);
int launch_target_f0n0_all(int nt, int ntp, const int* tperm, const double* d0,
                           const double* n0_orig, double* out, hipStream_t s);

}  // namespace hor
