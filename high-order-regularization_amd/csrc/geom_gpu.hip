// GPU geometry: the geometry-only inputs of the two-surface solve.
//
//   * quadrature (Generate_Surface, highreg_stokes_2surf/surface.cpp:13-71):
//     one thread per grid line, a count pass, a prefix sum and a write pass,
//     so the points come out in the reference's (plane, l1, l2, j) order;
//   * nearest-point data (Compute_nearest_data, surface.cpp:450-526): the
//     nearest quadrature point of the target's Monge plane through a uniform
//     cell grid with exact pruning, then Newton on (lambda, x);
//   * 6x6 Monge stencils (interpolate_DL_density + sixPoint_stencil,
//     main.cpp:682-748, 916-1046): candidates through tables of grid lines
//     instead of the reference's scan of every point, accepted by the
//     reference's coordinate tests.
//
// The level-set arithmetic is geom_core.h, the code the host geometry runs,
// compiled here without FMA contraction: positions, normals, nearest points,
// stencils and weights are bit-identical to the host's; only the
// partition-of-unity factor (exp, acos of the device math library) may
// differ from the host's in its last bits.
#pragma clang fp contract(off)

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "device.h"
#include "geom_core.h"
#include "geom_gpu.h"

namespace hor {
namespace {

constexpr int kScanThreads = 1024;

inline dim3 blocks_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// Exclusive prefix sum of n counts: out[i] = sum in[0..i), out[n] = total.
// One workgroup; thread t owns a contiguous chunk.
__global__ __launch_bounds__(kScanThreads) void k_scan(int n, const int* in, int* out) {
  __shared__ int part[kScanThreads];
  const int t = threadIdx.x;
  const int per = (n + kScanThreads - 1) / kScanThreads;
  const int b = min(n, t * per), e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int w = 1; w < kScanThreads; w <<= 1) {
    const int v = t >= w ? part[t - w] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; i++) {
    out[i] = run;
    run += in[i];
  }
  if (t == kScanThreads - 1) out[n] = part[t];
}

int scan(hipStream_t st, int n, const int* in, int* out) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, n, in, out);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

// Bounding box of n points: box[0..2] = min, box[3..5] = max.  One workgroup.
__global__ __launch_bounds__(kBlock) void k_bbox(int n, const double* x, double* box) {
  __shared__ double sm[6][kBlock];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = threadIdx.x; i < n; i += kBlock)
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], x[3 * (size_t)i + k]);
      hi[k] = fmax(hi[k], x[3 * (size_t)i + k]);
    }
  for (int k = 0; k < 3; k++) {
    sm[k][threadIdx.x] = lo[k];
    sm[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) {
        sm[k][threadIdx.x] = fmin(sm[k][threadIdx.x], sm[k][threadIdx.x + w]);
        sm[3 + k][threadIdx.x] = fmax(sm[3 + k][threadIdx.x], sm[3 + k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) box[threadIdx.x] = sm[threadIdx.x][0];
}

int bbox(hipStream_t st, int n, const double* x_dev, double box[6]) {
  DBuf<double> d;
  HOR_TRY(d.alloc(6));
  hipLaunchKernelGGL(k_bbox, dim3(1), dim3(kBlock), 0, st, n, x_dev, d.p);
  HOR_HIP_TRY(hipGetLastError());
  HOR_HIP_TRY(hipMemcpyAsync(box, d.p, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  return HOR_OK;
}

// ---------------------------------------------------------------------------
// Quadrature.

struct SurfLines {
  hor_body B;
  double h, theta, ct;  // ct = cos(theta)
  int N;                // l1, l2, j in [-N, N]
  int nlines;           // 3 (2N+1)^2, ordered (plane i, l1, l2)
};

template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_surface_lines(SurfLines S, int* count, const int* offset,
                                                          double* X, double* Nr, double* Ar, int* Pl) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= S.nlines) return;
  const int M = 2 * S.N + 1;
  const int i = line / (M * M);
  const int rem = line - i * M * M;
  const int l1 = rem / M - S.N, l2 = rem % M - S.N;
  const double h = S.h;
  double A[3] = {0.0, 0.0, 0.0}, Bp[3] = {0.0, 0.0, 0.0};
  if (i == 0) { A[1] = l1 * h; A[2] = l2 * h; Bp[1] = l1 * h; Bp[2] = l2 * h; }
  if (i == 1) { A[0] = l1 * h; A[2] = l2 * h; Bp[0] = l1 * h; Bp[2] = l2 * h; }
  if (i == 2) { A[0] = l1 * h; A[1] = l2 * h; Bp[0] = l1 * h; Bp[1] = l2 * h; }
  int c = 0;
  const int base = WRITE ? offset[line] : 0;
  for (int j = -S.N; j <= S.N; j++) {
    A[i] = j * h;
    Bp[i] = A[i] + h;
    if (geom::sign_phi(S.B, A) * geom::sign_phi(S.B, Bp) < 0) {
      const double tau = geom::newton(S.B, A, h, i, 0, 1);
      double x[3] = {A[0], A[1], A[2]};
      x[i] += tau * h;
      double n[3];
      for (int k = 0; k < 3; k++) n[k] = geom::dphi(S.B, k, x, false);
      const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int k = 0; k < 3; k++) n[k] /= nn;
      if (fabs(n[i]) >= S.ct) {
        if (WRITE) {
          const size_t o = (size_t)(base + c);
          const double area = h * h * geom::part_unity(i, n, S.theta) / fabs(n[i]);
          for (int k = 0; k < 3; k++) {
            X[3 * o + k] = x[k] + S.B.ctr[k];
            Nr[3 * o + k] = n[k];
          }
          Ar[o] = area;
          Pl[o] = i;
        }
        c++;
      }
    }
  }
  if (!WRITE) count[line] = c;
}

// ---------------------------------------------------------------------------
// Nearest quadrature point of the target's Monge plane.

struct CellRec {
  double lo[3], hi[3];  // bounding box of the cell's points
  int start, count;     // range in the binned arrays
};

struct CellGrid {
  double lo[3];
  double inv;  // 1 / cell size
  int nc[3];
  int ncell;   // nc0 nc1 nc2; key = plane * ncell + cell
};

__device__ inline int cell_key(const CellGrid& G, int plane, const double* x) {
  int c[3];
  for (int k = 0; k < 3; k++) {
    int v = (int)floor((x[k] - G.lo[k]) * G.inv);
    c[k] = min(max(v, 0), G.nc[k] - 1);
  }
  return plane * G.ncell + (c[0] * G.nc[1] + c[1]) * G.nc[2] + c[2];
}

__global__ void k_bin_count(int n, const double* sx, const int* splane, CellGrid G, int* cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int p = splane[j];
  if (p < 0 || p > 2) return;
  atomicAdd(&cnt[cell_key(G, p, sx + 3 * (size_t)j)], 1);
}

__global__ void k_bin_fill(int n, const double* sx, const int* splane, CellGrid G, const int* start,
                           int* fill, int* sidx, double* spt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int p = splane[j];
  if (p < 0 || p > 2) return;
  const int key = cell_key(G, p, sx + 3 * (size_t)j);
  const int pos = start[key] + atomicAdd(&fill[key], 1);
  sidx[pos] = j;
  for (int k = 0; k < 3; k++) spt[3 * (size_t)pos + k] = sx[3 * (size_t)j + k];
}

__global__ void k_nonempty(int nkeys, const int* start, int* flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nkeys) flag[k] = start[k + 1] > start[k] ? 1 : 0;
}

__global__ void k_cells(int nkeys, const int* start, const int* cpos, const double* spt, CellRec* cells) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys || start[k + 1] == start[k]) return;
  CellRec R;
  R.start = start[k];
  R.count = start[k + 1] - start[k];
  for (int d = 0; d < 3; d++) { R.lo[d] = 1e300; R.hi[d] = -1e300; }
  for (int q = R.start; q < R.start + R.count; q++)
    for (int d = 0; d < 3; d++) {
      R.lo[d] = fmin(R.lo[d], spt[3 * (size_t)q + d]);
      R.hi[d] = fmax(R.hi[d], spt[3 * (size_t)q + d]);
    }
  cells[cpos[k]] = R;
}

struct NearArgs {
  hor_body B;
  double DEL;
  int nt, ncell;
  const double* tx;
  const double* sx;       // caller order (Newton start)
  const CellRec* cells;   // plane p: [cpos[p ncell], cpos[(p+1) ncell])
  const int* cpos;
  const double* spt;      // binned coordinates
  const int* sidx;        // binned -> caller index
  int* nrst;
  double *x0, *n0, *b;
  int* i123;
  double* f0;             // may be null
  int* fails;
};

// One target per lane.  For the lane's plane, U2 = min over cells of the
// squared distance to the farthest corner of the cell's box bounds the
// squared distance to the nearest point; every cell whose box is nearer than
// U2 (with a rounding margin) is scanned exactly, with the reference's tie
// rule (first index among equal distances).  Cells are visited in a
// wave-uniform order, so their records and points are scalar loads.
__global__ __launch_bounds__(kBlock) void k_nearest(NearArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < A.nt;
  double pt[3] = {0.0, 0.0, 0.0};
  int i1 = 0, i2 = 1, i3 = -1;
  if (valid) {
    for (int k = 0; k < 3; k++) pt[k] = A.tx[3 * (size_t)i + k];
    geom::monge_patch(A.B, pt, i1, i2, i3);
  }
  int best = -1;
  double bd = 100.0;  // min_dist (surface.cpp:374), compared with squared distances
  for (int p = 0; p < 3; p++) {
    const bool mine = valid && i3 == p;
    if (!__any(mine)) continue;
    const int c0 = A.cpos[p * A.ncell], c1 = A.cpos[(p + 1) * A.ncell];
    double U2 = 1e300;
    for (int c = c0; c < c1; c++) {
      const CellRec& R = A.cells[c];
      double s = 0.0;
      for (int k = 0; k < 3; k++) {
        const double g = fmax(pt[k] - R.lo[k], R.hi[k] - pt[k]);
        s += g * g;
      }
      U2 = fmin(U2, s);
    }
    const double lim = U2 + 1e-12 * U2 + 1e-24;
    for (int c = c0; c < c1; c++) {
      const CellRec& R = A.cells[c];
      double s = 0.0;
      for (int k = 0; k < 3; k++) {
        const double g = fmax(fmax(R.lo[k] - pt[k], pt[k] - R.hi[k]), 0.0);
        s += g * g;
      }
      const bool need = mine && s <= lim;
      if (!__any(need)) continue;
      for (int q = R.start; q < R.start + R.count; q++) {
        const double* sp = A.spt + 3 * (size_t)q;
        const double d = (pt[0] - sp[0]) * (pt[0] - sp[0]) + (pt[1] - sp[1]) * (pt[1] - sp[1]) +
                         (pt[2] - sp[2]) * (pt[2] - sp[2]);
        const int j = A.sidx[q];
        if (need && (d < bd || (d == bd && best >= 0 && j < best))) {
          bd = d;
          best = j;
        }
      }
    }
  }
  if (!valid) return;
  // Compute_nearest_data's per-target steps (surface.cpp:450-526)
  if (A.f0) A.f0[3 * (size_t)i] = A.f0[3 * (size_t)i + 1] = A.f0[3 * (size_t)i + 2] = 0.0;
  A.nrst[i] = best;
  if (best < 0) {
    atomicAdd(A.fails, 1);
    return;
  }
  double xx0[3], nn0[3];
  if (geom::nearest_point(A.B, pt, A.sx + 3 * (size_t)best, xx0, nn0) == -1) {
    atomicAdd(A.fails, 1);
    return;
  }
  for (int k = 0; k < 3; k++) {
    A.x0[3 * (size_t)i + k] = xx0[k];
    A.n0[3 * (size_t)i + k] = nn0[k];
  }
  A.i123[3 * (size_t)i] = i1;
  A.i123[3 * (size_t)i + 1] = i2;
  A.i123[3 * (size_t)i + 2] = i3;
  const double bi = (pt[0] - xx0[0]) * nn0[0] + (pt[1] - xx0[1]) * nn0[1] + (pt[2] - xx0[2]) * nn0[2];
  A.b[i] = bi;
  if (fabs(bi) / A.DEL > 8.0) {
    A.nrst[i] = 0;  // far sentinel (surface.cpp:500-503)
    return;
  }
  if (A.f0) geom::surface_force(A.B, xx0, nn0, A.f0 + 3 * (size_t)i);
}

// ---------------------------------------------------------------------------
// Stencils.

// Grid lines of the source surface, one table per orientation o (the normal
// axis of a Monge patch): every point is filed under the rounded coordinates
// (l_a, l_b) of its two in-plane axes a < b, whatever its own plane, as the
// reference's coordinate match would find it.
struct LineTable {
  double ctr[3];
  double h;
  int lmin[3], lext[3];  // per axis, lines l in [lmin, lmin + lext)
  int toff[4];           // orientation o's table: [toff[o], toff[o+1])
};

__device__ inline void plane_axes(int o, int& a, int& b) {
  a = (o == 0) ? 1 : 0;
  b = (o == 2) ? 1 : 2;
}

__device__ inline int line_key(const LineTable& T, int o, long la, long lb) {
  int a, b;
  plane_axes(o, a, b);
  const long ia = la - T.lmin[a], ib = lb - T.lmin[b];
  if (ia < 0 || ia >= T.lext[a] || ib < 0 || ib >= T.lext[b]) return -1;
  return T.toff[o] + (int)ia * T.lext[b] + (int)ib;
}

__device__ inline int point_line_key(const LineTable& T, int o, const double* x) {
  int a, b;
  plane_axes(o, a, b);
  return line_key(T, o, lround((x[a] - T.ctr[a]) / T.h), lround((x[b] - T.ctr[b]) / T.h));
}

__global__ void k_line_count(int n, const double* sx, LineTable T, int* cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int o = 0; o < 3; o++) {
    const int key = point_line_key(T, o, sx + 3 * (size_t)j);
    if (key >= 0) atomicAdd(&cnt[key], 1);
  }
}

__global__ void k_line_fill(int n, const double* sx, LineTable T, const int* start, int* fill, int* lidx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int o = 0; o < 3; o++) {
    const int key = point_line_key(T, o, sx + 3 * (size_t)j);
    if (key >= 0) lidx[start[key] + atomicAdd(&fill[key], 1)] = j;
  }
}

struct StencilArgs {
  LineTable T;
  int nt;
  const double* sx;
  const int *nrst, *i123;
  const double* x0;
  const int *lstart, *lidx;
  int* st;
  double* co;
  int* bad;
};

// Slot [c][r] takes the largest source index whose in-plane coordinates equal
// the slot's node within 1e-14, first match in node order, on the same side
// (sign of the normal coordinate) as the target's nearest point; 0 when none
// (the reference's unset value).  A point that passes the tests rounds to the
// node's line, so the line's points are the only candidates.
__global__ __launch_bounds__(kBlock) void k_stencil(StencilArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nt) return;
  int* st = A.st + 36 * (size_t)i;
  double* co = A.co + 12 * (size_t)i;
  const int ind1 = A.i123[3 * (size_t)i], ind2 = A.i123[3 * (size_t)i + 1], ind3 = A.i123[3 * (size_t)i + 2];
  const bool ok = ind1 >= 0 && ind1 < 3 && ind2 >= 0 && ind2 < 3 && ind3 >= 0 && ind3 < 3 && ind1 != ind2 &&
                  ind1 != ind3 && ind2 != ind3;
  if (A.nrst[i] == 0 || !ok) {
    for (int s = 0; s < 36; s++) st[s] = 0;
    for (int s = 0; s < 12; s++) co[s] = 0.0;
    if (A.nrst[i] != 0) atomicAdd(A.bad, 1);
    return;
  }
  const double h = A.T.h;
  const double* p0 = A.x0 + 3 * (size_t)i;
  const double c1 = A.T.ctr[ind1], c2 = A.T.ctr[ind2], c3 = A.T.ctr[ind3];
  const double xLow = floor((p0[ind1] - c1) / h) * h + c1;
  const double yLow = floor((p0[ind2] - c2) / h) * h + c2;
  const double xHigh = xLow + h, yHigh = yLow + h;
  const double z0 = p0[ind3] - c3;
  const double cx[6] = {xLow - 2.0 * h, xLow - h, xLow, xHigh, xHigh + h, xHigh + 2.0 * h};
  const double cy[6] = {yLow - 2.0 * h, yLow - h, yLow, yHigh, yHigh + h, yHigh + 2.0 * h};
  const double tol = 1.e-14;
  for (int c = 0; c < 6; c++)
    for (int r = 0; r < 6; r++) {
      const long l1 = lround((cx[c] - c1) / h), l2 = lround((cy[r] - c2) / h);
      const int key = ind1 < ind2 ? line_key(A.T, ind3, l1, l2) : line_key(A.T, ind3, l2, l1);
      const int q0 = key >= 0 ? A.lstart[key] : 0, q1 = key >= 0 ? A.lstart[key + 1] : 0;
      int best = 0;
      for (int q = q0; q < q1; q++) {
        const int j = A.lidx[q];
        const double* s = A.sx + 3 * (size_t)j;
        const double z = s[ind3] - c3;
        if (!(((z0 > 0) && (z > 0)) || ((z0 < 0) && (z < 0)))) continue;
        const double X = s[ind1], Y = s[ind2];
        int col = -1, row = -1;
        for (int k = 0; k < 6; k++)
          if (fabs(cx[k] - X) < tol) { col = k; break; }
        if (col != c) continue;
        for (int k = 0; k < 6; k++)
          if (fabs(cy[k] - Y) < tol) { row = k; break; }
        if (row != r) continue;
        best = max(best, j);
      }
      st[6 * c + r] = best;
    }
  geom::six_coeff((p0[ind1] - xLow) / h, co);
  geom::six_coeff((p0[ind2] - yLow) / h, co + 6);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host drivers.

int surface_generate_gpu(hor_ctx* ctx, const hor_body& B, double h, double pou_deg, int cap, int* n,
                         double* x, double* nrml, double* area, int* plane) {
  hipStream_t st = ctx->stream;
  SurfLines S;
  S.B = B;
  S.h = h;
  const double rmax = std::max(B.abc[0], std::max(B.abc[1], B.abc[2]));
  S.N = (int)std::ceil(rmax / h) + 1;
  S.theta = pou_deg * kPi / 180.0;
  S.ct = std::cos(S.theta);
  const long M = 2L * S.N + 1;
  if (3 * M * M > INT_MAX / 2) return set_error(HOR_E_ARG, "surface generation: grid of %ld lines too large", 3 * M * M);
  S.nlines = (int)(3 * M * M);
  DBuf<int> cnt, off;
  HOR_TRY(cnt.alloc(S.nlines));
  HOR_TRY(off.alloc((size_t)S.nlines + 1));
  hipLaunchKernelGGL(k_surface_lines<false>, blocks_for(S.nlines), dim3(kBlock), 0, st, S, cnt.p, nullptr,
                     nullptr, nullptr, nullptr, nullptr);
  HOR_HIP_TRY(hipGetLastError());
  HOR_TRY(scan(st, S.nlines, cnt.p, off.p));
  int total = 0;
  HOR_HIP_TRY(hipMemcpyAsync(&total, off.p + S.nlines, sizeof(int), hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  *n = total;
  if (cap < total || total == 0) return HOR_OK;
  if (!x || !nrml || !area || !plane) return set_error(HOR_E_ARG, "hor_surface_generate_gpu: null output");
  DBuf<double> X, Nr, Ar;
  DBuf<int> Pl;
  HOR_TRY(X.alloc(3 * (size_t)total));
  HOR_TRY(Nr.alloc(3 * (size_t)total));
  HOR_TRY(Ar.alloc(total));
  HOR_TRY(Pl.alloc(total));
  hipLaunchKernelGGL(k_surface_lines<true>, blocks_for(S.nlines), dim3(kBlock), 0, st, S, nullptr, off.p, X.p,
                     Nr.p, Ar.p, Pl.p);
  HOR_HIP_TRY(hipGetLastError());
  HOR_HIP_TRY(hipMemcpyAsync(x, X.p, 3 * sizeof(double) * total, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(nrml, Nr.p, 3 * sizeof(double) * total, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(area, Ar.p, sizeof(double) * total, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(plane, Pl.p, sizeof(int) * total, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  return HOR_OK;
}

int nearest_data_gpu(hor_ctx* ctx, const hor_body& B, double h, double DEL, int ns, const double* sx,
                     const int* splane, int nt, const double* tx, int* nrst, double* x0, double* n0, double* b,
                     int* i123, double* f0) {
  hipStream_t st = ctx->stream;
  if (nt == 0) return HOR_OK;
  DBuf<double> dsx, dtx;
  DBuf<int> dpl;
  HOR_TRY(dsx.upload(sx, 3 * (size_t)ns, st));
  HOR_TRY(dpl.upload(splane, ns, st));
  HOR_TRY(dtx.upload(tx, 3 * (size_t)nt, st));
  double box[6];
  HOR_TRY(bbox(st, ns, dsx.p, box));
  // cells of about 8h, at most 2^21 per plane
  CellGrid G;
  double csize = 8.0 * h;
  for (;;) {
    long tot = 1;
    for (int k = 0; k < 3; k++) {
      G.nc[k] = std::max(1, (int)std::ceil((box[3 + k] - box[k]) / csize));
      tot *= G.nc[k];
    }
    if (tot <= (1L << 21)) break;
    csize *= 2.0;
  }
  for (int k = 0; k < 3; k++) G.lo[k] = box[k];
  G.inv = 1.0 / csize;
  G.ncell = G.nc[0] * G.nc[1] * G.nc[2];
  const int nkeys = 3 * G.ncell;
  DBuf<int> cnt, start, fill, flag, cpos, sidx, fails;
  DBuf<double> spt;
  DBuf<CellRec> cells;
  HOR_TRY(cnt.alloc(nkeys));
  HOR_TRY(fill.alloc(nkeys));
  HOR_TRY(flag.alloc(nkeys));
  HOR_TRY(start.alloc((size_t)nkeys + 1));
  HOR_TRY(cpos.alloc((size_t)nkeys + 1));
  HOR_TRY(sidx.alloc(ns));
  HOR_TRY(spt.alloc(3 * (size_t)ns));
  HOR_TRY(fails.alloc(1));
  HOR_HIP_TRY(hipMemsetAsync(cnt.p, 0, sizeof(int) * nkeys, st));
  HOR_HIP_TRY(hipMemsetAsync(fill.p, 0, sizeof(int) * nkeys, st));
  HOR_HIP_TRY(hipMemsetAsync(fails.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_bin_count, blocks_for(ns), dim3(kBlock), 0, st, ns, dsx.p, dpl.p, G, cnt.p);
  HOR_TRY(scan(st, nkeys, cnt.p, start.p));
  hipLaunchKernelGGL(k_bin_fill, blocks_for(ns), dim3(kBlock), 0, st, ns, dsx.p, dpl.p, G, start.p, fill.p,
                     sidx.p, spt.p);
  hipLaunchKernelGGL(k_nonempty, blocks_for(nkeys), dim3(kBlock), 0, st, nkeys, start.p, flag.p);
  HOR_TRY(scan(st, nkeys, flag.p, cpos.p));
  int ncells = 0;
  HOR_HIP_TRY(hipMemcpyAsync(&ncells, cpos.p + nkeys, sizeof(int), hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  HOR_TRY(cells.alloc(ncells));
  hipLaunchKernelGGL(k_cells, blocks_for(nkeys), dim3(kBlock), 0, st, nkeys, start.p, cpos.p, spt.p, cells.p);
  HOR_HIP_TRY(hipGetLastError());
  DBuf<int> dnrst, di123;
  DBuf<double> dx0, dn0, db, df0;
  HOR_TRY(dnrst.alloc(nt));
  HOR_TRY(di123.alloc(3 * (size_t)nt));
  HOR_TRY(dx0.alloc(3 * (size_t)nt));
  HOR_TRY(dn0.alloc(3 * (size_t)nt));
  HOR_TRY(db.alloc(nt));
  if (f0) HOR_TRY(df0.alloc(3 * (size_t)nt));
  NearArgs A;
  A.B = B;
  A.DEL = DEL;
  A.nt = nt;
  A.ncell = G.ncell;
  A.tx = dtx.p;
  A.sx = dsx.p;
  A.cells = cells.p;
  A.cpos = cpos.p;
  A.spt = spt.p;
  A.sidx = sidx.p;
  A.nrst = dnrst.p;
  A.x0 = dx0.p;
  A.n0 = dn0.p;
  A.b = db.p;
  A.i123 = di123.p;
  A.f0 = f0 ? df0.p : nullptr;
  A.fails = fails.p;
  hipLaunchKernelGGL(k_nearest, blocks_for(nt), dim3(kBlock), 0, st, A);
  HOR_HIP_TRY(hipGetLastError());
  int nfail = 0;
  HOR_HIP_TRY(hipMemcpyAsync(&nfail, fails.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(nrst, dnrst.p, sizeof(int) * nt, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(i123, di123.p, 3 * sizeof(int) * nt, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(x0, dx0.p, 3 * sizeof(double) * nt, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(n0, dn0.p, 3 * sizeof(double) * nt, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipMemcpyAsync(b, db.p, sizeof(double) * nt, hipMemcpyDeviceToHost, st));
  if (f0) HOR_HIP_TRY(hipMemcpyAsync(f0, df0.p, 3 * sizeof(double) * nt, hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  if (nfail) return set_error(HOR_E_ARG, "hor_nearest_data_gpu: nearest point not found for %d targets", nfail);
  return HOR_OK;
}

int stencil_build(hor_ctx* ctx, const double ctr[3], double h, int ns, const double* sx, int nt, const int* nrst, const double* x0, const int* i123, int* st_dev, double* co_dev) {
  hipStream_t st = ctx->stream;
  if (nt == 0) return HOR_OK;
  DBuf<double> dsx, dx0;
  DBuf<int> dnrst, di123;
  HOR_TRY(dsx.upload(sx, 3 * (size_t)ns, st));
  HOR_TRY(dnrst.upload(nrst, nt, st));
  HOR_TRY(dx0.upload(x0, 3 * (size_t)nt, st));
  HOR_TRY(di123.upload(i123, 3 * (size_t)nt, st));
  double box[6];
  HOR_TRY(bbox(st, ns, dsx.p, box));
  LineTable T;
  for (int k = 0; k < 3; k++) {
    T.ctr[k] = ctr[k];
    const double a = std::floor((box[k] - ctr[k]) / h) - 1.0, e = std::ceil((box[3 + k] - ctr[k]) / h) + 1.0;
    if (!(e - a < 1e5)) return set_error(HOR_E_ARG, "stencil build: surface spans too many grid lines");
    T.lmin[k] = (int)a;
    T.lext[k] = (int)(e - a) + 1;
  }
  T.h = h;
  const long keys = (long)T.lext[1] * T.lext[2] + (long)T.lext[0] * T.lext[2] + (long)T.lext[0] * T.lext[1];
  if (keys > (1L << 28)) return set_error(HOR_E_ARG, "stencil build: %ld grid lines, too many", keys);
  T.toff[0] = 0;
  T.toff[1] = T.toff[0] + T.lext[1] * T.lext[2];
  T.toff[2] = T.toff[1] + T.lext[0] * T.lext[2];
  T.toff[3] = T.toff[2] + T.lext[0] * T.lext[1];
  const int nkeys = T.toff[3];
  DBuf<int> cnt, start, fill, lidx, bad;
  HOR_TRY(cnt.alloc(nkeys));
  HOR_TRY(fill.alloc(nkeys));
  HOR_TRY(start.alloc((size_t)nkeys + 1));
  HOR_TRY(lidx.alloc(3 * (size_t)ns));
  HOR_TRY(bad.alloc(1));
  HOR_HIP_TRY(hipMemsetAsync(cnt.p, 0, sizeof(int) * nkeys, st));
  HOR_HIP_TRY(hipMemsetAsync(fill.p, 0, sizeof(int) * nkeys, st));
  HOR_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_line_count, blocks_for(ns), dim3(kBlock), 0, st, ns, dsx.p, T, cnt.p);
  HOR_TRY(scan(st, nkeys, cnt.p, start.p));
  hipLaunchKernelGGL(k_line_fill, blocks_for(ns), dim3(kBlock), 0, st, ns, dsx.p, T, start.p, fill.p, lidx.p);
  StencilArgs A;
  A.T = T;
  A.nt = nt;
  A.sx = dsx.p;
  A.nrst = dnrst.p;
  A.i123 = di123.p;
  A.x0 = dx0.p;
  A.lstart = start.p;
  A.lidx = lidx.p;
  A.st = st_dev;
  A.co = co_dev;
  A.bad = bad.p;
  hipLaunchKernelGGL(k_stencil, blocks_for(nt), dim3(kBlock), 0, st, A);
  HOR_HIP_TRY(hipGetLastError());
  int nbad = 0;
  HOR_HIP_TRY(hipMemcpyAsync(&nbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HOR_HIP_TRY(hipStreamSynchronize(st));
  if (nbad) return set_error(HOR_E_ARG, "stencil build: %d near targets with an invalid Monge patch (i123)", nbad);
  return HOR_OK;
}

}  // namespace hor

// ---------------------------------------------------------------------------
// C ABI.

using namespace hor;

extern "C" int hor_surface_generate_gpu(hor_ctx* ctx, const hor_body* body, double h, double pou_deg, int cap,
                                        int* n, double* x, double* nrml, double* area, int* plane) {
  if (!ctx || !body || !n || !(h > 0)) return set_error(HOR_E_ARG, "hor_surface_generate_gpu: bad argument");
  const double rmax = std::max(body->abc[0], std::max(body->abc[1], body->abc[2]));
  if (!(rmax > 0) || !(rmax / h < 1e6) || !(pou_deg > 0 && pou_deg <= 90))
    return set_error(HOR_E_ARG, "hor_surface_generate_gpu: need semi-axes > 0, h above max(abc)/1e6, 0 < pou_deg <= 90");
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  return surface_generate_gpu(ctx, *body, h, pou_deg, cap, n, x, nrml, area, plane);
}

extern "C" int hor_nearest_data_gpu(hor_ctx* ctx, const hor_body* body, double h, double DEL, int ns,
                                    const double* sx, const int* splane, int nt, const double* tx, int* nrst,
                                    double* x0, double* n0, double* b, int* i123, double* f0) {
  if (!ctx || !body || !sx || !splane || !tx || !nrst || !x0 || !n0 || !b || !i123 || ns <= 0 || nt < 0 ||
      !(h > 0) || !(DEL > 0))
    return set_error(HOR_E_ARG, "hor_nearest_data_gpu: bad argument");
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  return nearest_data_gpu(ctx, *body, h, DEL, ns, sx, splane, nt, tx, nrst, x0, n0, b, i123, f0);
}

extern "C" int hor_stencil_build(hor_ctx* ctx, const double* ctr, double h, int ns, const double* sx, int nt,
                                 const int* nrst, const double* x0, const int* i123, int* stencil,
                                 double* coeff) {
  if (!ctx || !ctr || !sx || !nrst || !x0 || !i123 || !stencil || !coeff || ns <= 0 || nt < 0 ||
      !(h > 0))
    return set_error(HOR_E_ARG, "hor_stencil_build: bad argument");
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  if (nt == 0) return HOR_OK;
  DBuf<int> dst;
  DBuf<double> dco;
  HOR_TRY(dst.alloc(36 * (size_t)nt));
  HOR_TRY(dco.alloc(12 * (size_t)nt));
  HOR_TRY(stencil_build(ctx, ctr, h, ns, sx, nt, nrst, x0, i123, dst.p, dco.p));
  HOR_HIP_TRY(hipMemcpyAsync(stencil, dst.p, 36 * sizeof(int) * nt, hipMemcpyDeviceToHost, ctx->stream));
  HOR_HIP_TRY(hipMemcpyAsync(coeff, dco.p, 12 * sizeof(double) * nt, hipMemcpyDeviceToHost, ctx->stream));
  HOR_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HOR_OK;
}
