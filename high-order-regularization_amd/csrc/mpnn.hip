// Learned right preconditioner of the two-surface GMRES: a point-cloud
// message-passing network z = M(r) over the quadrature points of both
// surfaces, with hand-written forward and backward kernels (BF16 matrix-core
// GEMMs with FP32 accumulation, FP32 elementwise and graph kernels) and an
// Adam step, so that it is trained on the GPU from (correction, residual)
// pairs logged by the solver's own GMRES.  It has no counterpart in the
// reference, whose GMRES is unpreconditioned (main.cpp:313-415).
//
// Network (width W = 64, k nearest neighbours over all points, L = 3 layers):
//   s      = rms_i |r_i| (per vector),  rh_i = r_i / s
//   x_in_i = [rh_i, n_i, p_i]           (p_i: position about the own body's
//                                        centre over the body's size)
//   h^0    = relu(W_in x_in + b_in)
//   layer l:  a = W_a h,  m_i = (1/k) sum_{j in N(i)} relu(a_j + B dl_ij + c),
//             dl_ij = (x_j - x_i) / rho_i  (rho_i: mean neighbour distance),
//             g_S = mean_{i in surface S} h_i,
//             h_i <- h_i + relu(W_u [h_i, m_i, g_S(i)] + d)
//   z_i    = r_i + s (W_out h^L_i + b_out)
// Training loss per pair (v, r = A v):  |z(r) - v|^2 / |v|^2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "mfma_gemm.h"
#include "plan.h"

namespace hor {

constexpr int kW = 64;
constexpr int kInPad = 16;  // input features, 9 used
constexpr int kLayers = 3;
constexpr int kUW = 3 * kW;  // node-update input width
// parameter layout (FP32, row-major weights [out][in])
constexpr int kOffWin = 0;
constexpr int kOffBin = kOffWin + kW * kInPad;
constexpr int kOffL0 = kOffBin + kW;
constexpr int kLWa = 0;
constexpr int kLBd = kLWa + kW * kW;     // [64][4], column 3 unused
constexpr int kLcc = kLBd + kW * 4;
constexpr int kLWu = kLcc + kW;
constexpr int kLd = kLWu + kW * kUW;
constexpr int kLSize = kLd + kW;
constexpr int kOffWout = kOffL0 + kLayers * kLSize;  // [3][64]
constexpr int kOffBout = kOffWout + 3 * kW;           // 3 (+1 pad)
constexpr int kNParams = kOffBout + 4;
constexpr int kChunk = 1024;  // rows per reduction chunk

// ---------------------------------------------------------------------------
// kernels

// Per-chunk sums of |r_i|^2 (rms scale) over the points of each vector.
__global__ void k_sq_partial(int npts, int nchunk_per, const double* r, double* part) {
  __shared__ double sh[256];
  const int b = blockIdx.y, ch = blockIdx.x;
  const int i0 = ch * kChunk, i1 = min(npts, i0 + kChunk);
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double* v = r + ((size_t)b * npts + i) * 3;
    s += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)b * nchunk_per + ch] = sh[0];
}

// out[b] = sum of the vector's chunk partials (in order); mode 0: rms scale
// sqrt(sum/npts), mode 1: plain sum.
__global__ void k_sq_final(int nb, int nchunk_per, int npts, const double* part, double* out, int mode) {
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nchunk_per; c++) s += part[(size_t)b * nchunk_per + c];
    out[b] = mode == 0 ? sqrt(s / npts) : s;
  }
}

// x_in = [r/s, n, p, 0...]
__global__ void k_features(int rows, int npts, const double* r, const double* scale, const float* geo,
                           float* xin) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const int b = row / npts, i = row - b * npts;
  const double s = scale[b];
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  float* x = xin + (size_t)row * kInPad;
  for (int c = 0; c < 3; c++) x[c] = (float)(r[(size_t)row * 3 + c] * inv);
  for (int c = 0; c < 6; c++) x[3 + c] = geo[(size_t)i * 8 + c];
  for (int c = 9; c < kInPad; c++) x[c] = 0.0f;
}

// pre += bias[c]; h = relu(pre)  (input layer), or h = hin + relu(pre) (residual)
__global__ void k_bias_act(int rows, const float* bias, float* pre, const float* hin, float* hout) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * kW) return;
  const int c = (int)(e % kW);
  const float p = pre[e] + bias[c];
  pre[e] = p;
  const float a = fmaxf(p, 0.0f);
  hout[e] = hin ? hin[e] + a : a;
}

// Segment partial column sums: chunk q covers rows [row0[q], row1[q]) of a
// rows x C matrix (leading dimension ld, column offset col0); part[q][c].
__global__ void k_colsum_partial(int C, int ld, int col0, const int* row0, const int* row1, const float* X,
                                 float* part) {
  const int q = blockIdx.x, c = threadIdx.x;
  if (c >= C) return;
  float s = 0.0f;
  for (int r = row0[q]; r < row1[q]; r++) s += X[(size_t)r * ld + col0 + c];
  part[(size_t)q * C + c] = s;
}

// out[seg][c] (+)= scale[seg] * sum over the segment's chunks (in order).
__global__ void k_colsum_final(int C, const int* seg_first, const int* seg_n, const float* part, const float* seg_scale,
                               float* out, int accumulate) {
  const int g = blockIdx.x, c = threadIdx.x;
  if (c >= C) return;
  float s = 0.0f;
  for (int q = seg_first[g]; q < seg_first[g] + seg_n[g]; q++) s += part[(size_t)q * C + c];
  if (seg_scale) s *= seg_scale[g];
  if (accumulate) out[(size_t)g * C + c] += s;
  else out[(size_t)g * C + c] = s;
}

// U = [h, m, g_seg]; m filled by k_edge_fwd.
__global__ void k_build_u(int rows, int npts, int n1, const float* h, const float* g, float* U) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * kW) return;
  const int row = (int)(e / kW), c = (int)(e % kW);
  const int b = row / npts, i = row - b * npts;
  const int seg = 2 * b + (i >= n1 ? 1 : 0);
  U[(size_t)row * kUW + c] = h[e];
  U[(size_t)row * kUW + 2 * kW + c] = g[(size_t)seg * kW + c];
}

// m_i = (1/k) sum_a relu(a_j + Bd dl_ia + cc), written to U[:, 64:128].
__global__ void k_edge_fwd(int rows, int npts, int k, const int* nbr, const float* dl, const float* Am,
                           const float* prm, float* U) {
  const int row = blockIdx.x * (blockDim.x / kW) + threadIdx.x / kW;
  const int c = threadIdx.x % kW;
  if (row >= rows) return;
  const int b = row / npts, i = row - b * npts;
  const size_t base = (size_t)b * npts;
  const float* Bd = prm + kLBd;
  const float b0 = Bd[c * 4], b1 = Bd[c * 4 + 1], b2 = Bd[c * 4 + 2], cc = prm[kLcc + c];
  float s = 0.0f;
  for (int a = 0; a < k; a++) {
    const int e = i * k + a;
    const int j = nbr[e];
    const float* d = dl + (size_t)e * 4;
    const float p = Am[(base + j) * kW + c] + b0 * d[0] + b1 * d[1] + b2 * d[2] + cc;
    s += fmaxf(p, 0.0f);
  }
  U[(size_t)row * kUW + kW + c] = s / (float)k;
}

// Backward of the edge messages through the reverse graph (deterministic, no
// atomics): dA_j = sum over edges (i -> j) of relu'(pre_ij) dm_i / k, and the
// per-block partial sums of the gradients of Bd and cc.
__global__ void k_edge_bwd(int rows, int npts, int k, const int* rev_off, const int* rev_edge, const float* dl,
                           const float* Am, const float* prm, const float* dU, float* dA, float* part) {
  __shared__ float sh[256 * 4];
  const int c = threadIdx.x % kW;
  const int slot = threadIdx.x / kW;  // 4 rows per pass
  const float* Bd = prm + kLBd;
  const float b0 = Bd[c * 4], b1 = Bd[c * 4 + 1], b2 = Bd[c * 4 + 2], cc = prm[kLcc + c];
  const float invk = 1.0f / (float)k;
  float gB0 = 0.f, gB1 = 0.f, gB2 = 0.f, gc = 0.f;
  for (int row = blockIdx.x * 4 + slot; row < rows; row += gridDim.x * 4) {
    const int b = row / npts, j = row - b * npts;
    const size_t base = (size_t)b * npts;
    const float aj = Am[(size_t)row * kW + c];
    float s = 0.0f;
    for (int q = rev_off[j]; q < rev_off[j + 1]; q++) {
      const int e = rev_edge[q];
      const int i = e / k;
      const float* d = dl + (size_t)e * 4;
      const float p = aj + b0 * d[0] + b1 * d[1] + b2 * d[2] + cc;
      if (p > 0.0f) {
        const float gm = dU[(base + i) * kUW + kW + c] * invk;
        s += gm;
        gB0 += gm * d[0];
        gB1 += gm * d[1];
        gB2 += gm * d[2];
        gc += gm;
      }
    }
    dA[(size_t)row * kW + c] = s;
  }
  sh[threadIdx.x * 4 + 0] = gB0;
  sh[threadIdx.x * 4 + 1] = gB1;
  sh[threadIdx.x * 4 + 2] = gB2;
  sh[threadIdx.x * 4 + 3] = gc;
  __syncthreads();
  if (slot == 0) {
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < 4; q++)
      for (int v = 0; v < 4; v++) t[v] += sh[(q * kW + c) * 4 + v];
    for (int v = 0; v < 4; v++) part[((size_t)blockIdx.x * kW + c) * 4 + v] = t[v];
  }
}

// gB[c][v] = sum over blocks of part (fixed order), v = 0..2 -> Bd grad, 3 -> cc grad
__global__ void k_edge_bwd_final(int nblk, const float* part, float* gBd, float* gcc) {
  const int c = threadIdx.x;
  if (c >= kW) return;
  float t[4] = {0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < nblk; q++)
    for (int v = 0; v < 4; v++) t[v] += part[((size_t)q * kW + c) * 4 + v];
  gBd[c * 4 + 0] = t[0];
  gBd[c * 4 + 1] = t[1];
  gBd[c * 4 + 2] = t[2];
  gBd[c * 4 + 3] = 0.0f;
  gcc[c] = t[3];
}

// z = r + s (o + b_out), FP64 out
__global__ void k_output(int rows, const double* r, const double* scale, int npts, const float* o,
                         const float* bout, double* z) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const int b = row / npts;
  const double s = scale[b];
  for (int c = 0; c < 3; c++)
    z[(size_t)row * 3 + c] = r[(size_t)row * 3 + c] + s * ((double)o[(size_t)row * 4 + c] + (double)bout[c]);
}

// out[c] = sum of the nq chunk partials (in order)
__global__ void k_colsum_total(int C, int nq, const float* part, float* out) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float s = 0.0f;
  for (int q = 0; q < nq; q++) s += part[(size_t)q * C + c];
  out[c] = s;
}

// Loss mean_b |z_b - v_b|^2/|v_b|^2: per-row contributions and the gradient
// with respect to the output-layer result o, dO = s 2 (z - v) / (|v_b|^2 batch)
// (FP32, leading dimension 4).
__global__ void k_loss_grad(int rows, int npts, int batch, const double* z, const double* v, const double* vnorm2,
                            const double* scale, float* dO, double* lrow) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const int b = row / npts;
  const double w = vnorm2[b] > 0.0 ? 1.0 / (vnorm2[b] * batch) : 0.0;
  double l = 0.0;
  for (int c = 0; c < 3; c++) {
    const double d = z[(size_t)row * 3 + c] - v[(size_t)row * 3 + c];
    l += d * d * w;
    dO[(size_t)row * 4 + c] = (float)(2.0 * d * w * scale[b]);
  }
  dO[(size_t)row * 4 + 3] = 0.0f;
  lrow[row] = l;
}

// dpre = dh * (pre > 0)
__global__ void k_relu_bwd(size_t n, const float* dh, const float* pre, float* dpre) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dpre[e] = pre[e] > 0.0f ? dh[e] : 0.0f;
}

// dh += dU[:, 0:64] + dg[seg] / n_seg
__global__ void k_dh_from_u(int rows, int npts, int n1, const float* dU, const float* dg, float* dh) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * kW) return;
  const int row = (int)(e / kW), c = (int)(e % kW);
  const int b = row / npts, i = row - b * npts;
  const int seg = 2 * b + (i >= n1 ? 1 : 0);
  const float ns = (float)(i >= n1 ? npts - n1 : n1);
  dh[e] += dU[(size_t)row * kUW + c] + dg[(size_t)seg * kW + c] / ns;
}

__global__ void k_sum_partials(int n, int nsplit, size_t stride, const float* part, float* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int z = 0; z < nsplit; z++) s += part[(size_t)z * stride + e];
  out[e] = s;
}

__global__ void k_adam(int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_sum_d(int n, const double* x, double* out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) s += x[e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

static inline dim3 g1(size_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace hor

using namespace hor;

struct hor_mpnn {
  hor_ctx* ctx = nullptr;
  int k = 16;
  DBuf<float> prm, grad, am, av;
  int adam_t = 0;
  // geometry
  int npts = 0, n1 = 0, batch_max = 0;
  DBuf<int> nbr, rev_off, rev_edge;
  DBuf<float> dl, geo;
  // reductions: chunks of the per-vector point range and of the (vector, surface) segments
  int nchunk_pt = 0;                // chunks per vector (sq sums)
  std::vector<int> seg_r0_h, seg_r1_h, seg_first_h, seg_n_h;
  DBuf<int> seg_r0, seg_r1, seg_first, seg_n;   // for batch_max vectors
  int nseg_chunks_per_b = 0;
  // activations (rows = batch * npts)
  int rows_cap = 0, rows = 0, batch = 0;
  DBuf<float> xin, pre0, H, Am, U, pre, o, gseg;
  DBuf<double> scale, vnorm2, sqpart, lrow, lsum, zbuf;
  // gradients
  DBuf<float> dO, dh, dU, dA, dpre, dg, part, epart;
  size_t part_cap = 0;
  const double* last_r = nullptr;
};

namespace {

constexpr int kEdgeBlocks = 1024;

int ensure_rows(hor_mpnn* N, int batch) {
  const int rows = batch * N->npts;
  if (rows <= N->rows_cap) return HOR_OK;
  HOR_HIP_TRY(hipStreamSynchronize(N->ctx->stream));  // buffers are replaced below
  const size_t R = rows;
  HOR_TRY(N->xin.alloc(R * kInPad));
  HOR_TRY(N->pre0.alloc(R * kW));
  HOR_TRY(N->H.alloc(R * kW * (kLayers + 1)));
  HOR_TRY(N->Am.alloc(R * kW * kLayers));
  HOR_TRY(N->U.alloc(R * kUW * kLayers));
  HOR_TRY(N->pre.alloc(R * kW * kLayers));
  HOR_TRY(N->o.alloc(R * 4));
  HOR_TRY(N->lrow.alloc(R));
  HOR_TRY(N->dO.alloc(R * 4));
  HOR_TRY(N->dh.alloc(R * kW));
  HOR_TRY(N->dU.alloc(R * kUW));
  HOR_TRY(N->dA.alloc(R * kW));
  HOR_TRY(N->dpre.alloc(R * kW));
  // partials: split-K products (at most 256 x 64 x 192) and chunk column sums
  const size_t nq = (size_t)batch * N->nseg_chunks_per_b;
  HOR_TRY(N->part.alloc(std::max((size_t)256 * kW * kUW, nq * kUW)));
  N->part_cap = N->part.n;
  N->rows_cap = rows;
  return HOR_OK;
}

// Partial buffers are sized by ensure_rows; a request beyond them is a bug.
int ensure_part(hor_mpnn* N, size_t n) {
  if (n <= N->part_cap) return HOR_OK;
  return set_error(HOR_E_STATE, "mpnn: partial buffer too small (%zu > %zu)", n, N->part_cap);
}

// C = op(A) op(B) with the BF16 matrix-core kernel; split-K (deterministic
// partials then an ordered sum) when K is long and the output small.
template <bool TA, bool TB>
int gemm(hor_mpnn* N, int M, int Nn, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
         int accumulate) {
  hipStream_t s = N->ctx->stream;
  const int tiles = ((M + 31) / 32) * ((Nn + 31) / 32);
  int nsplit = 1;
  if (tiles < 512 && K > 4096) nsplit = std::min(256, (K + 2047) / 2048);
  if (nsplit == 1) {
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), dim3((M + 31) / 32, (Nn + 31) / 32, 1), dim3(64), 0, s, M, Nn, K,
                       ((K + 7) / 8) * 8, A, lda, B, ldb, C, ldc, (size_t)0, accumulate);
    HOR_HIP_TRY(hipGetLastError());
    return HOR_OK;
  }
  if (accumulate || ldc != Nn) return set_error(HOR_E_ARG, "mpnn gemm: split-K output must be dense, not accumulated");
  const int kchunk = (((K + nsplit - 1) / nsplit) + 7) / 8 * 8;
  const int nz = (K + kchunk - 1) / kchunk;
  HOR_TRY(ensure_part(N, (size_t)nz * M * Nn));
  hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), dim3((M + 31) / 32, (Nn + 31) / 32, nz), dim3(64), 0, s, M, Nn, K, kchunk,
                     A, lda, B, ldb, N->part.p, Nn, (size_t)M * Nn, 0);
  hipLaunchKernelGGL(k_sum_partials, g1((size_t)M * Nn), dim3(256), 0, s, M * Nn, nz, (size_t)M * Nn, N->part.p, C);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

// column sums over all rows of a rows x C block -> out[C]
int colsum_all(hor_mpnn* N, int C, int ld, int col0, const float* X, float* out) {
  hipStream_t s = N->ctx->stream;
  const int nb = N->batch;
  const int nq = nb * N->nseg_chunks_per_b;
  HOR_TRY(ensure_part(N, (size_t)nq * C));
  hipLaunchKernelGGL(k_colsum_partial, dim3(nq), dim3(256), 0, s, C, ld, col0, N->seg_r0.p, N->seg_r1.p, X, N->part.p);
  hipLaunchKernelGGL(k_colsum_total, dim3(1), dim3(256), 0, s, C, nq, N->part.p, out);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

// per-(vector, surface) sums of columns [col0, col0+64) -> out[2*batch][64],
// scaled by 1/n_seg when mean
int segsum(hor_mpnn* N, int ld, int col0, const float* X, float* out, bool mean) {
  hipStream_t s = N->ctx->stream;
  const int nq = N->batch * N->nseg_chunks_per_b;
  HOR_TRY(ensure_part(N, (size_t)nq * kW));
  hipLaunchKernelGGL(k_colsum_partial, dim3(nq), dim3(256), 0, s, kW, ld, col0, N->seg_r0.p, N->seg_r1.p, X, N->part.p);
  hipLaunchKernelGGL(k_colsum_final, dim3(2 * N->batch), dim3(256), 0, s, kW, N->seg_first.p, N->seg_n.p, N->part.p,
                     mean ? N->epart.p : nullptr, out, 0);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

int forward(hor_mpnn* N, int batch, const double* r, double* z) {
  hipStream_t s = N->ctx->stream;
  if (batch < 1 || batch > N->batch_max) return set_error(HOR_E_ARG, "mpnn: batch %d outside [1, %d]", batch, N->batch_max);
  HOR_TRY(ensure_rows(N, batch));
  const int rows = batch * N->npts;
  N->rows = rows;
  N->batch = batch;
  const float* P = N->prm.p;
  // rms scale of each vector
  hipLaunchKernelGGL(k_sq_partial, dim3(N->nchunk_pt, batch), dim3(256), 0, s, N->npts, N->nchunk_pt, r, N->sqpart.p);
  hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(256), 0, s, batch, N->nchunk_pt, N->npts, N->sqpart.p, N->scale.p, 0);
  hipLaunchKernelGGL(k_features, g1(rows), dim3(256), 0, s, rows, N->npts, r, N->scale.p, N->geo.p, N->xin.p);
  HOR_HIP_TRY(hipGetLastError());
  // input layer
  HOR_TRY((gemm<false, true>(N, rows, kW, kInPad, N->xin.p, kInPad, P + kOffWin, kInPad, N->pre0.p, kW, 0)));
  float* H0 = N->H.p;
  hipLaunchKernelGGL(k_bias_act, g1((size_t)rows * kW), dim3(256), 0, s, rows, P + kOffBin, N->pre0.p, nullptr, H0);
  for (int l = 0; l < kLayers; l++) {
    const float* L = P + kOffL0 + (size_t)l * kLSize;
    float* Hl = N->H.p + (size_t)l * rows * kW;
    float* Hn = N->H.p + (size_t)(l + 1) * rows * kW;
    float* Al = N->Am.p + (size_t)l * rows * kW;
    float* Ul = N->U.p + (size_t)l * rows * kUW;
    float* Pl = N->pre.p + (size_t)l * rows * kW;
    HOR_TRY((gemm<false, true>(N, rows, kW, kW, Hl, kW, L + kLWa, kW, Al, kW, 0)));
    HOR_TRY(segsum(N, kW, 0, Hl, N->gseg.p + (size_t)l * 2 * N->batch_max * kW, true));
    hipLaunchKernelGGL(k_build_u, g1((size_t)rows * kW), dim3(256), 0, s, rows, N->npts, N->n1, Hl,
                       N->gseg.p + (size_t)l * 2 * N->batch_max * kW, Ul);
    hipLaunchKernelGGL(k_edge_fwd, g1((size_t)rows * kW), dim3(256), 0, s, rows, N->npts, N->k, N->nbr.p, N->dl.p, Al,
                       L, Ul);
    HOR_HIP_TRY(hipGetLastError());
    HOR_TRY((gemm<false, true>(N, rows, kW, kUW, Ul, kUW, L + kLWu, kUW, Pl, kW, 0)));
    hipLaunchKernelGGL(k_bias_act, g1((size_t)rows * kW), dim3(256), 0, s, rows, L + kLd, Pl, Hl, Hn);
  }
  const float* HL = N->H.p + (size_t)kLayers * rows * kW;
  HOR_TRY((gemm<false, true>(N, rows, 3, kW, HL, kW, P + kOffWout, kW, N->o.p, 4, 0)));
  hipLaunchKernelGGL(k_output, g1(rows), dim3(256), 0, s, rows, r, N->scale.p, N->npts, N->o.p, P + kOffBout, z);
  HOR_HIP_TRY(hipGetLastError());
  N->last_r = r;
  return HOR_OK;
}

int backward(hor_mpnn* N, const double* z, const double* v, double* loss) {
  hipStream_t s = N->ctx->stream;
  const int rows = N->rows, batch = N->batch;
  const float* P = N->prm.p;
  float* G = N->grad.p;
  // |v_b|^2
  hipLaunchKernelGGL(k_sq_partial, dim3(N->nchunk_pt, batch), dim3(256), 0, s, N->npts, N->nchunk_pt, v, N->sqpart.p);
  hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(256), 0, s, batch, N->nchunk_pt, N->npts, N->sqpart.p, N->vnorm2.p, 1);
  hipLaunchKernelGGL(k_loss_grad, g1(rows), dim3(256), 0, s, rows, N->npts, batch, z, v, N->vnorm2.p, N->scale.p,
                     N->dO.p, N->lrow.p);
  hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, s, rows, N->lrow.p, N->lsum.p);
  HOR_HIP_TRY(hipGetLastError());
  // output layer
  const float* HL = N->H.p + (size_t)kLayers * rows * kW;
  HOR_TRY(colsum_all(N, 4, 4, 0, N->dO.p, G + kOffBout));
  HOR_TRY((gemm<true, false>(N, 3, kW, rows, N->dO.p, 4, HL, kW, G + kOffWout, kW, 0)));
  HOR_TRY((gemm<false, false>(N, rows, kW, 3, N->dO.p, 4, P + kOffWout, kW, N->dh.p, kW, 0)));
  for (int l = kLayers - 1; l >= 0; l--) {
    const float* L = P + kOffL0 + (size_t)l * kLSize;
    float* GL = G + kOffL0 + (size_t)l * kLSize;
    const float* Hl = N->H.p + (size_t)l * rows * kW;
    const float* Al = N->Am.p + (size_t)l * rows * kW;
    const float* Ul = N->U.p + (size_t)l * rows * kUW;
    const float* Pl = N->pre.p + (size_t)l * rows * kW;
    hipLaunchKernelGGL(k_relu_bwd, g1((size_t)rows * kW), dim3(256), 0, s, (size_t)rows * kW, N->dh.p, Pl, N->dpre.p);
    HOR_TRY(colsum_all(N, kW, kW, 0, N->dpre.p, GL + kLd));
    HOR_TRY((gemm<true, false>(N, kW, kUW, rows, N->dpre.p, kW, Ul, kUW, GL + kLWu, kUW, 0)));
    HOR_TRY((gemm<false, false>(N, rows, kUW, kW, N->dpre.p, kW, L + kLWu, kUW, N->dU.p, kUW, 0)));
    // pooled context: dg[seg] = sum over the segment of dU[:, 128:192]
    DBuf<float>& dg = N->dg;
    HOR_TRY(segsum(N, kUW, 2 * kW, N->dU.p, dg.p, false));
    hipLaunchKernelGGL(k_dh_from_u, g1((size_t)rows * kW), dim3(256), 0, s, rows, N->npts, N->n1, N->dU.p, dg.p, N->dh.p);
    // edge messages
    const int nblk = std::min(kEdgeBlocks, (rows + 3) / 4);
    hipLaunchKernelGGL(k_edge_bwd, dim3(nblk), dim3(256), 0, s, rows, N->npts, N->k, N->rev_off.p, N->rev_edge.p,
                       N->dl.p, Al, L, N->dU.p, N->dA.p, N->epart.p + 2 * (size_t)N->batch_max);
    hipLaunchKernelGGL(k_edge_bwd_final, dim3(1), dim3(kW), 0, s, nblk, N->epart.p + 2 * (size_t)N->batch_max,
                       GL + kLBd, GL + kLcc);
    HOR_HIP_TRY(hipGetLastError());
    HOR_TRY((gemm<true, false>(N, kW, kW, rows, N->dA.p, kW, Hl, kW, GL + kLWa, kW, 0)));
    HOR_TRY((gemm<false, false>(N, rows, kW, kW, N->dA.p, kW, L + kLWa, kW, N->dh.p, kW, 1)));
  }
  hipLaunchKernelGGL(k_relu_bwd, g1((size_t)rows * kW), dim3(256), 0, s, (size_t)rows * kW, N->dh.p, N->pre0.p, N->dpre.p);
  HOR_TRY(colsum_all(N, kW, kW, 0, N->dpre.p, G + kOffBin));
  HOR_TRY((gemm<true, false>(N, kW, kInPad, rows, N->dpre.p, kW, N->xin.p, kInPad, G + kOffWin, kInPad, 0)));
  HOR_HIP_TRY(hipGetLastError());
  if (loss) {
    double l;
    HOR_HIP_TRY(hipMemcpyAsync(&l, N->lsum.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HOR_HIP_TRY(hipStreamSynchronize(s));
    *loss = l;  // the row terms carry the 1/batch of the mean
  }
  return HOR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" int hor_mpnn_param_count(void) { return kNParams; }

extern "C" int hor_mpnn_create(hor_ctx* ctx, int k, hor_mpnn** out) {
  if (!ctx || !out || k < 1 || k > 64) return set_error(HOR_E_ARG, "hor_mpnn_create: bad argument");
  *out = nullptr;
  HOR_HIP_TRY(hipSetDevice(ctx->device));
  std::unique_ptr<hor_mpnn> N(new hor_mpnn());
  N->ctx = ctx;
  N->k = k;
  HOR_TRY(N->prm.alloc(kNParams));
  HOR_TRY(N->grad.alloc(kNParams));
  HOR_TRY(N->am.alloc(kNParams));
  HOR_TRY(N->av.alloc(kNParams));
  HOR_HIP_TRY(hipMemsetAsync(N->prm.p, 0, sizeof(float) * kNParams, ctx->stream));
  HOR_HIP_TRY(hipMemsetAsync(N->grad.p, 0, sizeof(float) * kNParams, ctx->stream));
  HOR_HIP_TRY(hipMemsetAsync(N->am.p, 0, sizeof(float) * kNParams, ctx->stream));
  HOR_HIP_TRY(hipMemsetAsync(N->av.p, 0, sizeof(float) * kNParams, ctx->stream));
  HOR_HIP_TRY(hipStreamSynchronize(ctx->stream));
  *out = N.release();
  return HOR_OK;
}

extern "C" int hor_mpnn_destroy(hor_mpnn* N) {
  if (!N) return HOR_OK;
  (void)hipSetDevice(N->ctx->device);
  (void)hipStreamSynchronize(N->ctx->stream);
  delete N;
  return HOR_OK;
}

extern "C" int hor_mpnn_set_params(hor_mpnn* N, const float* p) {
  if (!N || !p) return set_error(HOR_E_ARG, "hor_mpnn_set_params: null");
  for (int e = 0; e < kNParams; e++)
    if (!std::isfinite(p[e])) return set_error(HOR_E_NAN, "hor_mpnn_set_params: non-finite parameter %d", e);
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  HOR_HIP_TRY(hipMemcpyAsync(N->prm.p, p, sizeof(float) * kNParams, hipMemcpyHostToDevice, N->ctx->stream));
  HOR_HIP_TRY(hipMemsetAsync(N->am.p, 0, sizeof(float) * kNParams, N->ctx->stream));
  HOR_HIP_TRY(hipMemsetAsync(N->av.p, 0, sizeof(float) * kNParams, N->ctx->stream));
  HOR_HIP_TRY(hipStreamSynchronize(N->ctx->stream));
  N->adam_t = 0;
  return HOR_OK;
}

extern "C" int hor_mpnn_get_params(hor_mpnn* N, float* p) {
  if (!N || !p) return set_error(HOR_E_ARG, "hor_mpnn_get_params: null");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  HOR_HIP_TRY(hipMemcpyAsync(p, N->prm.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, N->ctx->stream));
  HOR_HIP_TRY(hipStreamSynchronize(N->ctx->stream));
  return HOR_OK;
}

extern "C" int hor_mpnn_get_grad(hor_mpnn* N, float* g) {
  if (!N || !g) return set_error(HOR_E_ARG, "hor_mpnn_get_grad: null");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  HOR_HIP_TRY(hipMemcpyAsync(g, N->grad.p, sizeof(float) * kNParams, hipMemcpyDeviceToHost, N->ctx->stream));
  HOR_HIP_TRY(hipStreamSynchronize(N->ctx->stream));
  return HOR_OK;
}

// Point set of the two surfaces in the solver's unknown order: n1 points of
// surface 1 then n2 of surface 2; ctr (6) and size (2) per body.
extern "C" int hor_mpnn_set_geometry(hor_mpnn* N, int n1, int n2, const double* x, const double* nrml,
                                     const double* ctr, const double* size, int batch_max, int* nbr_out) {
  if (!N || n1 < 1 || n2 < 1 || !x || !nrml || !ctr || !size || batch_max < 1)
    return set_error(HOR_E_ARG, "hor_mpnn_set_geometry: bad argument");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  hipStream_t s = N->ctx->stream;
  HOR_HIP_TRY(hipStreamSynchronize(s));  // the graph buffers are replaced below
  const int npts = n1 + n2, k = N->k;
  if ((long long)npts * batch_max > (1LL << 30)) return set_error(HOR_E_ARG, "hor_mpnn_set_geometry: too many rows");
  std::vector<int> nbr;
  HOR_TRY(knn_graph(npts, x, k, nbr));
  if (nbr_out) std::memcpy(nbr_out, nbr.data(), sizeof(int) * nbr.size());
  // edge offsets dl = (x_j - x_i)/rho_i, rho_i = mean neighbour distance
  std::vector<float> dl((size_t)npts * k * 4, 0.0f), geo((size_t)npts * 8, 0.0f);
  for (int i = 0; i < npts; i++) {
    double rho = 0.0;
    for (int a = 0; a < k; a++) {
      const int j = nbr[(size_t)i * k + a];
      double d2 = 0.0;
      for (int c = 0; c < 3; c++) d2 += (x[3 * j + c] - x[3 * i + c]) * (x[3 * j + c] - x[3 * i + c]);
      rho += std::sqrt(d2);
    }
    rho = rho > 0.0 ? rho / k : 1.0;
    for (int a = 0; a < k; a++) {
      const int j = nbr[(size_t)i * k + a];
      for (int c = 0; c < 3; c++) dl[((size_t)i * k + a) * 4 + c] = (float)((x[3 * j + c] - x[3 * i + c]) / rho);
    }
    const int body = i < n1 ? 0 : 1;
    for (int c = 0; c < 3; c++) {
      geo[(size_t)i * 8 + c] = (float)nrml[3 * i + c];
      geo[(size_t)i * 8 + 3 + c] = (float)((x[3 * i + c] - ctr[3 * body + c]) / size[body]);
    }
  }
  // reverse graph: edges (i, a) grouped by their neighbour j, in (i, a) order
  std::vector<int> roff(npts + 1, 0), redge((size_t)npts * k);
  for (size_t e = 0; e < nbr.size(); e++) roff[nbr[e] + 1]++;
  for (int j = 0; j < npts; j++) roff[j + 1] += roff[j];
  std::vector<int> fill(roff.begin(), roff.end() - 1);
  for (size_t e = 0; e < nbr.size(); e++) redge[fill[nbr[e]]++] = (int)e;
  HOR_TRY(N->nbr.upload(nbr, s));
  HOR_TRY(N->rev_off.upload(roff, s));
  HOR_TRY(N->rev_edge.upload(redge, s));
  HOR_TRY(N->dl.upload(dl, s));
  HOR_TRY(N->geo.upload(geo, s));
  N->npts = npts;
  N->n1 = n1;
  N->batch_max = batch_max;
  N->nchunk_pt = (npts + kChunk - 1) / kChunk;
  // segment chunk tables: per vector b, surface 1 rows then surface 2 rows
  std::vector<int> r0, r1, sf, sn, ssc;
  std::vector<float> inv_n;
  for (int b = 0; b < batch_max; b++)
    for (int sg = 0; sg < 2; sg++) {
      const int a0 = b * npts + (sg == 0 ? 0 : n1), a1 = b * npts + (sg == 0 ? n1 : npts);
      sf.push_back((int)r0.size());
      for (int a = a0; a < a1; a += kChunk) {
        r0.push_back(a);
        r1.push_back(std::min(a1, a + kChunk));
      }
      sn.push_back((int)r0.size() - sf.back());
      inv_n.push_back(1.0f / (float)(a1 - a0));
    }
  N->nseg_chunks_per_b = (int)r0.size() / batch_max;
  HOR_TRY(N->seg_r0.upload(r0, s));
  HOR_TRY(N->seg_r1.upload(r1, s));
  HOR_TRY(N->seg_first.upload(sf, s));
  HOR_TRY(N->seg_n.upload(sn, s));
  // epart: [0, 2*batch_max) = 1/n_seg, then the edge-gradient partials
  std::vector<float> ep(2 * (size_t)batch_max + (size_t)kEdgeBlocks * kW * 4, 0.0f);
  for (size_t q = 0; q < inv_n.size(); q++) ep[q] = inv_n[q];
  HOR_TRY(N->epart.upload(ep, s));
  HOR_TRY(N->gseg.alloc((size_t)kLayers * 2 * batch_max * kW));
  HOR_TRY(N->dg.alloc(2 * (size_t)batch_max * kW));
  HOR_TRY(N->scale.alloc(batch_max));
  HOR_TRY(N->vnorm2.alloc(batch_max));
  HOR_TRY(N->sqpart.alloc((size_t)batch_max * N->nchunk_pt));
  HOR_TRY(N->lsum.alloc(1));
  HOR_TRY(N->zbuf.alloc((size_t)batch_max * npts * 3));
  N->rows_cap = 0;
  N->rows = 0;
  HOR_HIP_TRY(hipStreamSynchronize(s));
  return HOR_OK;
}

extern "C" int hor_mpnn_check_size(hor_mpnn* N, int n) {
  if (!N) return set_error(HOR_E_ARG, "hor_mpnn_check_size: null");
  if (N->npts == 0) return set_error(HOR_E_ARG, "mpnn: no geometry set");
  if (3 * N->npts != n) return set_error(HOR_E_ARG, "mpnn: geometry has %d unknowns, solver %d", 3 * N->npts, n);
  return HOR_OK;
}

// z = M(r) for batch vectors of 3*npts values (device pointers).
extern "C" int hor_mpnn_forward_dev(hor_mpnn* N, int batch, const double* r_dev, double* z_dev) {
  if (!N || !r_dev || !z_dev) return set_error(HOR_E_ARG, "hor_mpnn_forward_dev: null");
  if (N->npts == 0) return set_error(HOR_E_ARG, "mpnn: no geometry set");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  HOR_TRY(forward(N, batch, r_dev, z_dev));
  HOR_HIP_TRY(hipStreamSynchronize(N->ctx->stream));
  return HOR_OK;
}

// Preconditioner application inside the solver's stream (no synchronization).
extern "C" int hor_mpnn_apply_dev(hor_mpnn* N, hor_ctx* ctx, const double* r_dev, double* z_dev) {
  if (!N || !r_dev || !z_dev) return set_error(HOR_E_ARG, "hor_mpnn_apply_dev: null");
  if (ctx != N->ctx) return set_error(HOR_E_ARG, "hor_mpnn_apply_dev: network and solver use different contexts");
  return forward(N, 1, r_dev, z_dev);
}

// Forward on r then the gradient of mean_b |z_b - v_b|^2/|v_b|^2 with respect
// to the parameters (kept on the device; hor_mpnn_get_grad copies it out).
extern "C" int hor_mpnn_loss_grad_dev(hor_mpnn* N, int batch, const double* r_dev, const double* v_dev, double* loss) {
  if (!N || !r_dev || !v_dev) return set_error(HOR_E_ARG, "hor_mpnn_loss_grad_dev: null");
  if (N->npts == 0) return set_error(HOR_E_ARG, "mpnn: no geometry set");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  HOR_TRY(forward(N, batch, r_dev, N->zbuf.p));
  HOR_TRY(backward(N, N->zbuf.p, v_dev, loss));  // loss: the batch mean
  return HOR_OK;
}

extern "C" int hor_mpnn_adam_step(hor_mpnn* N, float lr, float beta1, float beta2, float eps) {
  if (!N || !(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
    return set_error(HOR_E_ARG, "hor_mpnn_adam_step: bad argument");
  HOR_HIP_TRY(hipSetDevice(N->ctx->device));
  hipStream_t s = N->ctx->stream;
  N->adam_t++;
  const float c1 = 1.0f - std::pow(beta1, (float)N->adam_t), c2 = 1.0f - std::pow(beta2, (float)N->adam_t);
  hipLaunchKernelGGL(k_adam, g1(kNParams), dim3(256), 0, s, kNParams, N->prm.p, N->grad.p, N->am.p, N->av.p, lr,
                     beta1, beta2, eps, c1, c2);
  HOR_HIP_TRY(hipGetLastError());
  HOR_HIP_TRY(hipStreamSynchronize(s));
  return HOR_OK;
}
