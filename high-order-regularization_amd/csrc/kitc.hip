// KITC treecode on gfx950 (KITC.h, KITC.cpp).
//
// Set-up (geometry only, once per plan): the octree over the sources and
// every target's interaction list are built on the host exactly as the
// reference's recursive traversal decides them; the lists are transposed into
// per-node work items (a node and up to 256 targets that visit it), and the
// barycentric factors a1, a2, a3 and D of every source in every far-field
// node are tabulated.
//
// Evaluation (per density, all on the device):
//   1. k_moments:  proxy weights W[kk][c] = sum_s (a1_i a2_j a3_k D)(s) (A q_c)(s)
//                  over the node's members, a GEMM on the FP64 matrix cores
//                  (proxies x sources times sources x components);
//   2. k_leaf:     regularized direct sums of a leaf's sources for its near
//                  targets (same pair code and LDS tiling as the direct kernel);
//   3. k_far:      plain Stokeslet / stresslet from the (P+1)^3 proxies with
//                  target-side subtraction (Call_Treecode_SL/DL);
//   4. k_tree_finalize: 16 lanes per target add its slots (a fixed-order
//                  reduction), scale, subtract and add the jump term.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pairs.h"
#include "plan.h"

namespace hor {


// Members per moment work item: enough items to give the moment GEMM about
// 2048 workgroups, between kMomChunkMin and kMomChunkMax members each.
constexpr int kMomChunkMin = 256, kMomChunkMax = 4096, kMomWantBlocks = 2048;
constexpr int kMomReduceBlock = 256;
// Near sums: split the leaves' sources when fewer than kNearWantItems items
// (workgroups) would run, into at most kNearMaxSplit chunks.  8192 items keep
// more waves resident through the on-surface near sums (h = 1/256 FP64 matvec
// 60.7 -> 58.2 ms, mixed 42.7 -> 40.2 ms against 2048).
constexpr int kNearWantItems = 8192, kNearMaxSplit = 64;

struct TreePlan {
  Tree T;
  TargetLists L;
  int np1 = 0, pflat = 0, ncomp = 0;
  int nslot = 0;
  // node data on the device
  DBuf<int> node_lo, node_hi, leaf_sub_off, far_idx;
  DBuf<int> slot_sub;        // per tree slot: the leaf subtile holding it (FP32 record centre)
  DBuf<double> cheb, leaf_box;
  // near work items
  int n_near_items = 0;
  DBuf<int> near_item_node, near_item_off, near_item_cnt, near_tgt, near_slot;
  DBuf<int> near_item_src;   // per near item: source slots [j0, j1), first subtile, chunk index
  // far work items
  int n_far_items = 0;
  DBuf<int> far_item_node, far_item_off, far_item_cnt, far_tgt, far_slot;
  // moments
  int nfar_nodes = 0, n_mom_items = 0;
  DBuf<int> far_node, fac_off, mom_item_far, mom_item_m0, mom_item_m1, mom_node_first, mom_node_nitems;
  DBuf<double> mom_part;     // per member-chunk work item: [kk][ncomp]
  DBuf<double> fac;          // per far node, member-major: a1 (np1), a2, a3, D per member
  DBuf<double> mq;           // per tree slot: A q_c, 16 values (zero-padded)
  DBuf<double> mom;          // per far node, proxy-major [kk][ncomp]
  DBuf<float> mom32;             // mixed precision: [kk][ncomp32] FP32 copy (padded)
  DBuf<float> cheb32;            // per node: Chebyshev points minus the node centre, FP32
  DBuf<double> node_ctr;         // per node: centre (midpoint of the Chebyshev interval)
  std::vector<char> far_flag;    // host, per node: moments needed (far for some target)
  // Set when another plan on the same source tree computes this plan's
  // moments (its far nodes included) and is evaluated first on the same
  // density: the far field reads that plan's far_idx, mom and mom32.
  const TreePlan* mom_owner = nullptr;
  // mixed-precision single-layer far field: HOR_FAR_VALU (k_far) or
  // HOR_FAR_MFMA (k_far_sl_mfma over the per-proxy columns in bmat)
  int far_kernel = HOR_FAR_VALU;
  DBuf<float> bmat;
  // reduction
  DBuf<int> slot_start;      // sorted target order, nt+1
  DBuf<double> slots;        // 3*nslot
  // Double layer: the target-side subtraction of every pair term, g - g0 with
  // g0 the target's density (KITC.cpp:1461, 1678), is linear in g0 with
  // coefficients that depend only on the geometry:
  //   sum_j K_ij (g_j - g0_i) = sum_j K_ij g_j - C_i g0_i,
  // C_i the sum of the near pairs' K_ij (the leaf sums subtract only at a near
  // target of the off-surface plan) and of the far proxies' (W_n.d) d d^T / r^5.
  // csub holds scale * C_i, SoA [(3 k + l) ntp + i] in sorted target order,
  // computed at plan creation by the plan's own FP64 kernels on zero density
  // with g0 = -e_l (the three columns); every later evaluation runs the
  // kernels on g alone (NOSUB) and k_tree_finalize subtracts csub g0.
  // A plan computes them when asked (tree_plan_prepare_subtraction, which the
  // solver calls for its plans before they share moments) or at its second
  // evaluation, so a plan evaluated once does not pay for them.
  DBuf<double> csub;
  bool has_csub = false;
  int n_evals = 0;
  // phase events of the last evaluation: before the moments, after the
  // moments, after the near sums, after the far field
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool evaluated = false;
  ~TreePlan() {
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// ---------------------------------------------------------------------------
// Device kernels

struct MomArgs {
  int np1, pflat, ncomp;
  const int *item_far, *item_m0, *item_m1;  // per work item: far node index, member range [m0, m1)
  const int *far_node, *node_lo, *fac_off;
  const double* fac;   // member-major barycentric factors (a1, a2, a3, D) per far node
  const double* mq;    // per tree slot: A q_c (16 values)
  double* part;        // per work item: [kk][ncomp] partial weights
};

// A q_c of every source (tree order): SL q = (f, n), DL q = (g_l n_m, n)
// (Panel_Moments_SL / Panel_Moments_DL, KITC.cpp:476-487, 632-651).
__global__ void k_moment_rhs(int ns, int nsp, int dl, const double* sn, const double* sa, const double* sd,
                             double* mq) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ns) return;
  const double a = sa[p];
  const double n[3] = {sn[p], sn[nsp + p], sn[2 * nsp + p]};
  const double d[3] = {sd[p], sd[nsp + p], sd[2 * nsp + p]};
  double* o = mq + (size_t)p * 16;
  if (dl) {
    // d^T W_{g n} d only sees the symmetric part of W_{g n}: diagonal, then
    // the three sums of the off-diagonal pairs, then W_n
    o[0] = a * d[0] * n[0];
    o[1] = a * d[1] * n[1];
    o[2] = a * d[2] * n[2];
    o[3] = a * (d[0] * n[1] + d[1] * n[0]);
    o[4] = a * (d[0] * n[2] + d[2] * n[0]);
    o[5] = a * (d[1] * n[2] + d[2] * n[1]);
    for (int m = 0; m < 3; m++) o[6 + m] = a * n[m];
    for (int c = 9; c < 16; c++) o[c] = 0.0;
  } else {
    for (int m = 0; m < 3; m++) {
      o[m] = a * d[m];
      o[3 + m] = a * n[m];
    }
    for (int c = 6; c < 16; c++) o[c] = 0.0;
  }
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Proxy weights (Panel_Moments_*, KITC.cpp:351-668) as a GEMM on the FP64
// matrix cores.  With kk = (i, j, k) the weights of component c factor as
//   W[(j,k)][(i,c)] = sum_s A[(j,k)][s] B[s][(i,c)],
//   A[(j,k)][s] = a2_j a3_k,   B[s][(i,c)] = a1_i D q_c   (q = A q of member s),
// a GEMM of np1^2 rows and np1*ncomp columns, every column a weight.  A work
// item is a far node's member chunk; each wave owns 16 rows and all column
// tiles (at most kMomTiles) and runs through the chunk four members per
// v_mfma_f64_16x16x4f64: lane l holds A[l%16][l/16] (one value per step,
// shared by the tiles) and B[l/16][16 t + l%16] for tile t; accumulator r of
// tile t in lane l is W[l/16 + 4 r][16 t + l%16].  The tiles' products are
// independent, so consecutive MFMAs do not wait on each other.
constexpr int kMomTiles = 8;   // np1*ncomp <= 13*9 = 117 columns (P <= 12)

// One step of four members for the NT column tiles (operands in LDS).
template <int NT>
__device__ __forceinline__ void moment_step(const double* fm, const double* qm, int aj, int ak, int dj,
                                            const int* bi, const int* bc, f64x4_t* acc) {
  const double a = fm[aj] * fm[ak];
  const double d = fm[dj];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const double b = fm[bi[t]] * d * qm[bc[t]];
    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
  }
}

// Members staged in LDS per round: their barycentric-factor rows (up to
// 3 * 13 + 1 values for P <= 12) and A q rows (16 values), loaded coalesced
// by the whole workgroup; the four waves read their operands from LDS.
constexpr int kMomStage = 32, kMomFacMax = 3 * 13 + 1;

template <int NT>
__global__ __launch_bounds__(kBlock) void k_moments(MomArgs A) {
  __shared__ double sfac[kMomStage * kMomFacMax];
  __shared__ double sq[kMomStage * 16];
  const int item = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int np1 = A.np1, ncomp = A.ncomp, nrow = np1 * np1, ncol = np1 * ncomp;
  const int row0 = (blockIdx.y * 4 + wave) * 16;
  const bool rows_ok = row0 < nrow;  // wave-uniform; every wave stages members and meets the barriers
  const int f = A.item_far[item];
  const int node = A.far_node[f];
  const int lo = A.node_lo[node];
  const int c0 = A.item_m0[item], c1 = A.item_m1[item];  // member range within the node
  const int stride = 3 * np1 + 1;
  const int ms = lane >> 4, l16 = lane & 15;
  // the lane's A row (j, k) and, per tile, its B column (i, c); rows and
  // columns past the ends use the last one and are not stored
  const int ar = min(row0 + l16, nrow - 1);
  const int aj = np1 + ar / np1, ak = 2 * np1 + ar % np1, dj = 3 * np1;
  int bi[NT], bc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = min(16 * t + l16, ncol - 1);
    bi[t] = col / ncomp;
    bc[t] = col - bi[t] * ncomp;
  }
  f64x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f64x4_t{0.0, 0.0, 0.0, 0.0};
  const double* fsrc = A.fac + A.fac_off[f] + (size_t)c0 * stride;
  const double* qsrc = A.mq + (size_t)(lo + c0) * 16;
  const int n = c1 - c0;
  for (int m = 0; m < n; m += kMomStage) {
    const int nm = min(kMomStage, n - m);
    __syncthreads();  // the previous round's operands have been read
    // members past the chunk are zero rows: their products add nothing
    for (int e = threadIdx.x; e < kMomStage * stride; e += kBlock)
      sfac[e] = e < nm * stride ? fsrc[(size_t)m * stride + e] : 0.0;
    for (int e = threadIdx.x; e < kMomStage * 16; e += kBlock) sq[e] = e < nm * 16 ? qsrc[(size_t)m * 16 + e] : 0.0;
    __syncthreads();
    if (rows_ok) {
      const int steps = (nm + 3) / 4;
      for (int st = 0; st < steps; st++)
        moment_step<NT>(sfac + (st * 4 + ms) * stride, sq + (st * 4 + ms) * 16, aj, ak, dj, bi, bc, acc);
    }
  }
  if (!rows_ok) return;
  double* P = A.part + (size_t)item * A.pflat * ncomp;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = 16 * t + l16;
    if (col >= ncol) continue;
    const int i = col / ncomp, c = col - i * ncomp;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = row0 + ms + 4 * r;   // j np1 + k
      if (row < nrow) P[((size_t)i * nrow + row) * ncomp + c] = acc[t][r];
    }
  }
}

// Sums the member-chunk partials of each far node in chunk order into the
// proxy-major moments (and their FP32 copy, padded to ncomp32 per proxy):
// one thread per (far node, moment value), blockIdx.x the node.
__global__ void k_moments_reduce(int ncomp, int pflat, const int* node_first, const int* node_nitems,
                                 const double* part, double* mom, float* mom32, int ncomp32) {
  const int f = blockIdx.x;
  const int total = ncomp * pflat;
  const int e = blockIdx.y * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int it0 = node_first[f], nit = node_nitems[f];
  const double* p = part + (size_t)it0 * total + e;
  double s = 0.0;
  for (int q = 0; q < nit; q++) s += p[(size_t)q * total];
  mom[(size_t)f * total + e] = s;
  if (mom32) {
    const int kk = e / ncomp, c = e - kk * ncomp;
    mom32[((size_t)f * pflat + kk) * ncomp32 + c] = (float)s;
  }
}

struct LeafArgs {
  DirectArgs D;
  const int *item_off, *item_cnt, *tgt, *slot;
  const int* item_src;   // 4 per item: source slots j0, j1, first leaf subtile, chunk (slot offset)
  double* slots;
  int nslot;
};

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Leaf direct sums (Call_Ds_*, KITC.cpp:807-1062) for up to 256 near targets
// over one chunk of a leaf's sources.  NOSUB: the double layer on g, not
// g - g0 (see TreePlan::csub).
template <int KIND, bool MIXED, bool NOSUB>
__global__ __launch_bounds__(kBlock) void k_leaf(LeafArgs A) {
  const double* tab = stage_table<KIND>(A.D);   // the workgroup's only barrier
  const int tid = threadIdx.x;
  const int item = blockIdx.x;
  const int off = A.item_off[item], cnt = A.item_cnt[item];
  // a wave without targets leaves at once (the last item of a node's list is
  // partial)
  if ((tid & ~(kWave - 1)) >= cnt) return;
  const bool act = tid < cnt;
  const int li = off + (act ? tid : 0);
  const int i = A.tgt[li];
  Tgt t;
  load_target<KIND, NOSUB>(A.D, i, t);
  // wave bounding box and class
  double wb[6];
  wb[0] = wave_min(t.px); wb[1] = wave_min(t.py); wb[2] = wave_min(t.pz);
  wb[3] = wave_max(t.px); wb[4] = wave_max(t.py); wb[5] = wave_max(t.pz);
  const bool off_kind = (KIND == HOR_SL_OFF || KIND == HOR_DL_OFF);
  const bool all_far = off_kind && !__any(t.nearp ? 1 : 0);
  double u[3] = {0.0, 0.0, 0.0};
  const int* src = A.item_src + 4 * item;   // this item's chunk of the leaf's sources
  range_sum<KIND, MIXED, MIXED>(A.D, src[0], src[1], src[2], wb, all_far, t, tab, u);
  if (act) {
    int s = A.slot[li] + src[3];
    A.slots[s] = u[0];
    A.slots[A.nslot + s] = u[1];
    A.slots[2 * A.nslot + s] = u[2];
  }
}

struct FarArgs {
  int np1, pflat, ntp, nslot;
  const int *item_node, *item_off, *item_cnt, *tgt, *slot, *far_idx;
  const double *cheb, *mom, *tx, *tsub2, *node_ctr;
  const float *cheb32, *mom32;
  double* slots;
};

constexpr int kNc32Sl = 8;    // FP32 moment floats per proxy (6 used)
constexpr int kNc32Dl = 12;   // (9 used)

// Call_Treecode_SL / Call_Treecode_DL (KITC.cpp:671-777): plain Stokeslet /
// stresslet from the (P+1)^3 proxies of one far node, with the target-side
// subtraction folded in:
//   SL: u += m / r + d (m.d) / r^3,             m = W_f - (f0.n0) W_n
//   DL: u += d [d.M.d - (g0.d)(W_n.d)] / r^5,   M_lm = W_{g_l n_m}
// (the DL form is the reference's sum_l d_l sum_m (W_lm - g0_l W_n,m) d_m;
// d.M.d uses only the symmetric part of M, so the moments hold its diagonal
// and the three off-diagonal sums: 9 components with W_n).
// One work item is a far node and up to 256 of its targets; the node's
// proxies and moments are wave-uniform (scalar loads).  Each lane carries T
// targets (target tid + q * 256/T of the item for q < T): every moment loaded
// serves T targets and the T chains of arithmetic interleave.  In mixed
// precision the proxy interactions run in FP32 on node-centred coordinates,
// each row of P+1 proxies is summed in FP32 and the rows in FP64.  NOSUB:
// the stresslet without its g0 term -(g0.d)(W_n.d) (see TreePlan::csub).
template <bool DL, bool MIXED, int T, bool NOSUB>
__global__ __launch_bounds__(kBlock / T) void k_far(FarArgs A) {
  constexpr int NT = kBlock / T;           // threads per workgroup
  const int tid = threadIdx.x;
  const int item = blockIdx.x;
  const int node = A.item_node[item];
  const int off = A.item_off[item], cnt = A.item_cnt[item];
  if ((tid & ~(kWave - 1)) >= cnt) return;  // a wave without targets (no workgroup barrier here)
  bool act[T];
  int li[T], ti[T];
#pragma unroll
  for (int q = 0; q < T; q++) {
    const int k = tid + q * NT;
    act[q] = k < cnt;
    li[q] = off + (act[q] ? k : 0);
    ti[q] = A.tgt[li[q]];
  }
  const int np1 = A.np1, pflat = A.pflat;
  double px[T], py[T], pz[T];
#pragma unroll
  for (int q = 0; q < T; q++) {
    px[q] = A.tx[ti[q]];
    py[q] = A.tx[A.ntp + ti[q]];
    pz[q] = A.tx[2 * A.ntp + ti[q]];
  }
  const int fi = A.far_idx[node];
  double u0[T], u1[T], u2[T];
#pragma unroll
  for (int q = 0; q < T; q++) u0[q] = u1[q] = u2[q] = 0.0;
  if (MIXED) {
    const float* t1 = A.cheb32 + (size_t)3 * np1 * node;
    const float* t2 = t1 + np1;
    const float* t3 = t2 + np1;
    const double* ctr = A.node_ctr + 3 * node;
    float qx[T], qy[T], qz[T];
#pragma unroll
    for (int q = 0; q < T; q++) {
      qx[q] = (float)(px[q] - ctr[0]);
      qy[q] = (float)(py[q] - ctr[1]);
      qz[q] = (float)(pz[q] - ctr[2]);
    }
    constexpr int NC = DL ? kNc32Dl : kNc32Sl;
    const float* M = A.mom32 + (size_t)fi * pflat * NC;
    if (!DL) {
      float sub[T];
#pragma unroll
      for (int q = 0; q < T; q++) sub[q] = (float)A.tsub2[ti[q]];
      for (int a = 0, kk = 0; a < np1; a++) {
        float d0[T];
#pragma unroll
        for (int q = 0; q < T; q++) d0[q] = qx[q] - t1[a];
        for (int b = 0; b < np1; b++) {
          float d1[T], r2ab[T], v0[T], v1[T], v2[T];
#pragma unroll
          for (int q = 0; q < T; q++) {
            d1[q] = qy[q] - t2[b];
            r2ab[q] = d0[q] * d0[q] + d1[q] * d1[q];
            v0[q] = v1[q] = v2[q] = 0.f;
          }
#pragma unroll 4
          for (int c = 0; c < np1; c++, kk++) {
            const float* m = M + (size_t)kk * NC;
#pragma unroll
            for (int q = 0; q < T; q++) {
              const float d2 = qz[q] - t3[c];
              const float ri = rsqrt32(__builtin_fmaf(d2, d2, r2ab[q]));
              const float m0 = m[0] - sub[q] * m[3], m1 = m[1] - sub[q] * m[4], m2 = m[2] - sub[q] * m[5];
              const float tq = (m0 * d0[q] + m1 * d1[q] + m2 * d2) * (ri * ri);
              v0[q] += (m0 + tq * d0[q]) * ri;
              v1[q] += (m1 + tq * d1[q]) * ri;
              v2[q] += (m2 + tq * d2) * ri;
            }
          }
#pragma unroll
          for (int q = 0; q < T; q++) {
            u0[q] += (double)v0[q]; u1[q] += (double)v1[q]; u2[q] += (double)v2[q];
          }
        }
      }
    } else {
      float g0x[T], g0y[T], g0z[T];
#pragma unroll
      for (int q = 0; q < T; q++) {
        g0x[q] = (float)A.tsub2[ti[q]];
        g0y[q] = (float)A.tsub2[A.ntp + ti[q]];
        g0z[q] = (float)A.tsub2[2 * A.ntp + ti[q]];
      }
      for (int a = 0, kk = 0; a < np1; a++) {
        float d0[T], d00[T];
#pragma unroll
        for (int q = 0; q < T; q++) {
          d0[q] = qx[q] - t1[a];
          d00[q] = d0[q] * d0[q];
        }
        for (int b = 0; b < np1; b++) {
          float d1[T], r2ab[T], gab[T], d11[T], d01[T], kab[T], v2[T];
#pragma unroll
          for (int q = 0; q < T; q++) {
            d1[q] = qy[q] - t2[b];
            r2ab[q] = __builtin_fmaf(d1[q], d1[q], d00[q]);
            gab[q] = g0x[q] * d0[q] + g0y[q] * d1[q];
            d11[q] = d1[q] * d1[q];
            d01[q] = d0[q] * d1[q];
            kab[q] = v2[q] = 0.f;
          }
#pragma unroll 4
          for (int c = 0; c < np1; c++, kk++) {
            const float* m = M + (size_t)kk * NC;
#pragma unroll
            for (int q = 0; q < T; q++) {
              const float d2 = qz[q] - t3[c];
              const float ri = rsqrt32(__builtin_fmaf(d2, d2, r2ab[q]));
              const float ri2 = ri * ri;
              const float qab = __builtin_fmaf(m[3], d01[q], __builtin_fmaf(m[1], d11[q], m[0] * d00[q]));
              const float w = __builtin_fmaf(m[2], d2, __builtin_fmaf(m[5], d1[q], m[4] * d0[q]));
              float Q = __builtin_fmaf(w, d2, qab);
              if constexpr (!NOSUB) {
                const float wnd = m[6] * d0[q] + m[7] * d1[q] + m[8] * d2;
                const float gd = __builtin_fmaf(g0z[q], d2, gab[q]);
                Q = __builtin_fmaf(-gd, wnd, Q);
              }
              const float kern = Q * (ri2 * ri2 * ri);
              kab[q] += kern;
              v2[q] += d2 * kern;
            }
          }
#pragma unroll
          for (int q = 0; q < T; q++) {
            u0[q] += (double)(d0[q] * kab[q]); u1[q] += (double)(d1[q] * kab[q]); u2[q] += (double)v2[q];
          }
        }
      }
    }
  } else {
    const double* t1 = A.cheb + (size_t)3 * np1 * node;
    const double* t2 = t1 + np1;
    const double* t3 = t2 + np1;
    constexpr int NC = DL ? 9 : 6;
    const double* M = A.mom + (size_t)fi * pflat * NC;
    if (!DL) {
      double sub[T];
#pragma unroll
      for (int q = 0; q < T; q++) sub[q] = A.tsub2[ti[q]];
      for (int a = 0, kk = 0; a < np1; a++) {
        double d0[T];
#pragma unroll
        for (int q = 0; q < T; q++) d0[q] = px[q] - t1[a];
        for (int b = 0; b < np1; b++) {
          double d1[T], r2ab[T];
#pragma unroll
          for (int q = 0; q < T; q++) {
            d1[q] = py[q] - t2[b];
            r2ab[q] = d0[q] * d0[q] + d1[q] * d1[q];
          }
#pragma unroll 2
          for (int c = 0; c < np1; c++, kk++) {
            const double* m = M + (size_t)kk * NC;
#pragma unroll
            for (int q = 0; q < T; q++) {
              const double d2 = pz[q] - t3[c];
              const double ri = rsqrt64(__builtin_fma(d2, d2, r2ab[q]));
              const double m0 = m[0] - sub[q] * m[3], m1 = m[1] - sub[q] * m[4], m2 = m[2] - sub[q] * m[5];
              const double tq = (m0 * d0[q] + m1 * d1[q] + m2 * d2) * (ri * ri);
              u0[q] += (m0 + tq * d0[q]) * ri;
              u1[q] += (m1 + tq * d1[q]) * ri;
              u2[q] += (m2 + tq * d2) * ri;
            }
          }
        }
      }
    } else {
      const double k5 = vreg(4.375);   // rsqrt5_64's e^2 coefficient, in vector registers for the loops
      double g0x[T], g0y[T], g0z[T];
#pragma unroll
      for (int q = 0; q < T; q++) {
        g0x[q] = A.tsub2[ti[q]];
        g0y[q] = A.tsub2[A.ntp + ti[q]];
        g0z[q] = A.tsub2[2 * A.ntp + ti[q]];
      }
      for (int a = 0, kk = 0; a < np1; a++) {
        double d0[T], d00[T];
#pragma unroll
        for (int q = 0; q < T; q++) {
          d0[q] = px[q] - t1[a];
          d00[q] = d0[q] * d0[q];
        }
        for (int b = 0; b < np1; b++) {
          double d1[T], r2ab[T], gab[T], d11[T], d01[T], kab[T];
#pragma unroll
          for (int q = 0; q < T; q++) {
            d1[q] = py[q] - t2[b];
            r2ab[q] = __builtin_fma(d1[q], d1[q], d00[q]);
            gab[q] = g0x[q] * d0[q] + g0y[q] * d1[q];
            d11[q] = d1[q] * d1[q];
            d01[q] = d0[q] * d1[q];
            kab[q] = 0.0;  // the row's kernel sum: d0 and d1 are fixed along it
          }
#pragma unroll 2
          for (int c = 0; c < np1; c++, kk++) {
            const double* m = M + (size_t)kk * NC;
#pragma unroll
            for (int q = 0; q < T; q++) {
              const double d2 = pz[q] - t3[c];
              const double ri5 = rsqrt5_64(__builtin_fma(d2, d2, r2ab[q]), k5);
              // d^T M d = M00 d0^2 + M11 d1^2 + S01 d0 d1 + d2 (S02 d0 + S12 d1 + M22 d2)
              const double qab = __builtin_fma(m[3], d01[q], __builtin_fma(m[1], d11[q], m[0] * d00[q]));
              const double w = __builtin_fma(m[2], d2, __builtin_fma(m[5], d1[q], m[4] * d0[q]));
              double Q = __builtin_fma(w, d2, qab);
              if constexpr (!NOSUB) {
                const double wnd = m[6] * d0[q] + m[7] * d1[q] + m[8] * d2;
                const double gd = __builtin_fma(g0z[q], d2, gab[q]);
                Q = __builtin_fma(-gd, wnd, Q);
              }
              const double kern = Q * ri5;
              kab[q] += kern;
              u2[q] += d2 * kern;
            }
          }
#pragma unroll
          for (int q = 0; q < T; q++) {
            u0[q] += d0[q] * kab[q];
            u1[q] += d1[q] * kab[q];
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < T; q++) {
    if (act[q]) {
      const int s = A.slot[li[q]];
      A.slots[s] = u0[q];
      A.slots[A.nslot + s] = u1[q];
      A.slots[2 * A.nslot + s] = u2[q];
    }
  }
}

// Targets per lane of the far-field kernel: two in mixed precision (h = 1/256
// mixed matvec 38.5 -> 38.0 ms), one in FP64, where two per lane halve the
// occupancy (120 VGPRs) for the same time.
constexpr int kFarT64 = 1, kFarT32 = 2;

// ---------------------------------------------------------------------------
// Single-layer far field as a GEMM on FP32 matrix cores (mixed precision,
// chosen per plan with hor_plan_set_far_kernel).  The proxy interaction of
// Call_Treecode_SL (KITC.cpp:671-720), w/r + d (w.d)/r^3 with d = p - y in
// node-centred coordinates (p the target, y the proxy), is K3 = 1/r^3 times
// r^2 w + d (w.d), a polynomial in p and y.  Per target the sums over the
// proxies of K3 times
//   S = w, T = w.y, U_ij = y_i w_j, V = y (w.y), Z = |y|^2 w
// (19 columns per moment set) give
//   u = |p|^2 S - 2 U^T p + Z + p (p.S) - p T - U p + V,
// and the subtraction's two moment sets (w = W_f, W_n) make 38 columns of the
// B operand, padded to 48, whose rows are the node's proxies.  The A operand
// [targets x proxies] is K3 from the VALU.  Each v_mfma_f32_16x16x4f32 takes
// 16 targets x 4 proxies x 16 columns; the FP32 products are summed over
// kFarMfmaFlush groups of four proxies on the matrix core and then added in
// FP64, and the per-target combination runs in FP64.
constexpr int kFarMfmaCols = 48;   // 2 x 19 moment columns, 2 zero, proxy position at 40..42, 5 zero
constexpr int kFarMfmaFlush = 8;   // 32 proxies per FP32 partial sum (k_far sums rows of P + 1)

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// B rows of every far node of the moment owner: [far node][proxy][48], the
// proxies padded to a multiple of four with zero moment columns (their
// position is the last proxy's, so K3 stays finite).
__global__ void k_far_sl_bmat(int nfar, int np1, int pflat, int pflat4, const int* far_node, const float* cheb32,
                              const float* mom32, float* B) {
  const int fi = blockIdx.y;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= nfar || kk >= pflat4) return;
  const int kc = min(kk, pflat - 1);
  const int a = kc / (np1 * np1), b = (kc / np1) % np1, c = kc % np1;
  const float* t = cheb32 + (size_t)3 * np1 * far_node[fi];
  const double y[3] = {t[a], t[np1 + b], t[2 * np1 + c]};
  const double yy = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  const float* m = mom32 + ((size_t)fi * pflat + kc) * kNc32Sl;
  float* o = B + ((size_t)fi * pflat4 + kk) * kFarMfmaCols;
  const bool live = kk < pflat;
  for (int s = 0; s < 2; s++) {
    double w[3];
    for (int j = 0; j < 3; j++) w[j] = live ? (double)m[3 * s + j] : 0.0;
    const double wy = w[0] * y[0] + w[1] * y[1] + w[2] * y[2];
    float* col = o + 19 * s;
    for (int j = 0; j < 3; j++) {
      col[j] = (float)w[j];
      col[13 + j] = (float)(y[j] * wy);
      col[16 + j] = (float)(yy * w[j]);
      for (int i = 0; i < 3; i++) col[4 + 3 * i + j] = (float)(y[i] * w[j]);
    }
    col[3] = (float)wy;
  }
  o[38] = o[39] = 0.0f;
  for (int j = 0; j < 3; j++) o[40 + j] = (float)y[j];
  for (int j = 43; j < kFarMfmaCols; j++) o[j] = 0.0f;
}

struct FarMfmaArgs {
  int pflat4, ntp, nslot;
  const int *item_node, *item_off, *item_cnt, *tgt, *slot, *far_idx;
  const double *tx, *tsub2, *node_ctr;
  const float* B;
  double* slots;
};

// u of one moment set from its 19 column sums (the formula above), FP64.
__device__ __forceinline__ void sl_far_combine(const double* v, const double p[3], double u[3]) {
  const double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  const double pS = p[0] * v[0] + p[1] * v[1] + p[2] * v[2];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double utp = 0.0, up = 0.0;   // (U^T p)_i and (U p)_i
#pragma unroll
    for (int k = 0; k < 3; k++) {
      utp += p[k] * v[4 + 3 * k + i];
      up += v[4 + 3 * i + k] * p[k];
    }
    u[i] = pp * v[i] - 2.0 * utp + v[16 + i] + p[i] * (pS - v[3]) - up + v[13 + i];
  }
}

// One work item is a far node and up to 256 of its targets (as k_far); wave
// w takes targets [64 w, 64 w + 64) as four tiles of 16 rows.  Lane l
// supplies the A value of target row l % 16 and proxy l / 16 of each group of
// four, and the B values of that proxy at column l % 16 of the three column
// tiles; result row 4 (l / 16) + r and column l % 16 of a tile are in
// element r of the lane's accumulator.
__global__ __launch_bounds__(kBlock) void k_far_sl_mfma(FarMfmaArgs A) {
  __shared__ double tile[kBlock / kWave][16][kFarMfmaCols + 1];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int item = blockIdx.x;
  const int node = A.item_node[item], off = A.item_off[item], cnt = A.item_cnt[item];
  const int w0 = wave * kWave;
  if (w0 >= cnt) return;   // a wave without targets (no workgroup barrier here)
  const double cx = A.node_ctr[3 * node], cy = A.node_ctr[3 * node + 1], cz = A.node_ctr[3 * node + 2];
  const int l16 = lane & 15, kq = lane >> 4;
  float qx[4], qy[4], qz[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int k = w0 + 16 * r + l16;
    const int ti = A.tgt[off + (k < cnt ? k : 0)];   // rows past the end repeat the item's first target
    qx[r] = (float)(A.tx[ti] - cx);
    qy[r] = (float)(A.tx[A.ntp + ti] - cy);
    qz[r] = (float)(A.tx[2 * A.ntp + ti] - cz);
  }
  const float* Bn = A.B + (size_t)A.far_idx[node] * A.pflat4 * kFarMfmaCols;
  f32x4_t acc[4][3];
  double s64[4][3][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[r][c] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int i = 0; i < 4; i++) s64[r][c][i] = 0.0;
    }
  const int ng = A.pflat4 / 4;
  for (int g0 = 0; g0 < ng; g0 += kFarMfmaFlush) {
    const int g1 = min(ng, g0 + kFarMfmaFlush);
    for (int g = g0; g < g1; g++) {
      const float* row = Bn + (size_t)(4 * g + kq) * kFarMfmaCols;
      const float b0 = row[l16], b1 = row[16 + l16], b2 = row[32 + l16];
      const float yx = row[40], yy = row[41], yz = row[42];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float dx = qx[r] - yx, dy = qy[r] - yy, dz = qz[r] - yz;
        const float ri = rsqrt32(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
        const float k3 = ri * ri * ri;
        acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(k3, b0, acc[r][0], 0, 0, 0);
        acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(k3, b1, acc[r][1], 0, 0, 0);
        acc[r][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(k3, b2, acc[r][2], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int i = 0; i < 4; i++) s64[r][c][i] += (double)acc[r][c][i];
        acc[r][c] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
      }
  }
  // per tile: column sums through this wave's LDS rows, then one lane per
  // target row combines them (LDS operations of one wave complete in order)
  double(*T)[kFarMfmaCols + 1] = tile[wave];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    if (w0 + 16 * r >= cnt) break;   // wave-uniform
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int i = 0; i < 4; i++) T[4 * kq + i][16 * c + l16] = s64[r][c][i];
    __builtin_amdgcn_wave_barrier();
    const int k = w0 + 16 * r + l16;
    if (kq == 0 && k < cnt) {
      const int li = off + k;
      const int ti = A.tgt[li];
      const double p[3] = {A.tx[ti] - cx, A.tx[A.ntp + ti] - cy, A.tx[2 * A.ntp + ti] - cz};
      double uf[3], un[3];
      sl_far_combine(T[l16], p, uf);
      sl_far_combine(T[l16] + 19, p, un);
      const double sub = A.tsub2[ti];
      const int s = A.slot[li];
      A.slots[s] = uf[0] - sub * un[0];
      A.slots[A.nslot + s] = uf[1] - sub * un[1];
      A.slots[2 * A.nslot + s] = uf[2] - sub * un[2];
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Target data of coefficient column l (TreePlan::csub): g0 = -e_l for the far
// field (tsub2) and for the leaf sums (tsub) at every target of the
// on-surface plan and at the near targets of the off-surface plan.
__global__ void k_csub_targets(int ntp, int l, int near_only, const int* tnear, double* tsub, double* tsub2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp) return;
  for (int k = 0; k < 3; k++) {
    const double v = k == l ? -1.0 : 0.0;
    tsub2[(size_t)k * ntp + i] = v;
    tsub[(size_t)k * ntp + i] = (!near_only || tnear[i]) ? v : 0.0;
  }
}

// csub[(3 k + l) ntp + i] <- col[3 tperm[i] + k], column l of the
// coefficients in sorted target order.
__global__ void k_csub_store(int nt, int ntp, int l, const int* tperm, const double* col, double* csub) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int o = tperm[i];
  for (int k = 0; k < 3; k++) csub[(size_t)(3 * k + l) * ntp + i] = col[3 * (size_t)o + k];
}

// Per-target reduction of the slots (16 lanes per target, four targets per
// wavefront): lane l of a target's group adds its slots a + l, a + l + 16, ...
// of each component (coalesced reads of the target's contiguous slot range),
// a butterfly adds the 16 partial sums (every lane of the group ends with the
// same total, in the same order on every run), and lane k < 3 finishes
// component k: v = scale*sum, minus the subtraction coefficients' C g0
// (NOSUB), plus the jump term xi*g0, combined into out as the direct finalize
// does.
constexpr int kFinLanes = 16;   // lanes per target
__global__ __launch_bounds__(kBlock) void k_tree_finalize(int nt, int ntp, int nslot, const int* tperm,
                                                          const int* slot_start, const double* slots, double scale,
                                                          const double* xi, const double* g0, const double* csub,
                                                          double* out, const double* base, double coef,
                                                          int accumulate) {
  const int lane = threadIdx.x & (kFinLanes - 1);
  const int i = blockIdx.x * (kBlock / kFinLanes) + threadIdx.x / kFinLanes;
  if (i >= nt) return;   // the target's whole lane group
  const int a = slot_start[i], b = slot_start[i + 1];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int q = a + lane; q < b; q += kFinLanes) {
    s0 += slots[q];
    s1 += slots[(size_t)nslot + q];
    s2 += slots[2 * (size_t)nslot + q];
  }
  for (int m = kFinLanes / 2; m > 0; m >>= 1) {
    s0 += __shfl_xor(s0, m);
    s1 += __shfl_xor(s1, m);
    s2 += __shfl_xor(s2, m);
  }
  if (lane >= 3) return;
  const int k = lane;   // component
  const int o = tperm[i];
  double v = (k == 0 ? s0 : (k == 1 ? s1 : s2)) * scale;
  if (csub) {
    const double* c = csub + (size_t)3 * k * ntp + i;
    v -= (c[0] * g0[i] + c[ntp] * g0[ntp + i]) + c[2 * (size_t)ntp] * g0[2 * (size_t)ntp + i];
  }
  if (xi) v += xi[i] * g0[(size_t)k * ntp + i];
  double w = coef * v;
  if (accumulate) out[3 * o + k] += w;
  else if (base) out[3 * o + k] = base[3 * o + k] + w;
  else out[3 * o + k] = w;
}

// ---------------------------------------------------------------------------
// Host set-up

// Barycentric factors of one source for one node's Chebyshev points
// (Panel_Moments_*, KITC.cpp:398-459): a_j = w_j/(y - t_j), snapped to the
// node when |y - t_j| < 1e-7, D = 1/(sum a1 sum a2 sum a3).
static void bary_factors(int P, const double* t1, const double* t2, const double* t3, double x, double y,
                         double z, double* a1, double* a2, double* a3, double* D) {
  int np1 = P + 1;
  double S[3] = {0, 0, 0};
  int flag[3] = {-1, -1, -1};
  const double* T[3] = {t1, t2, t3};
  double v[3] = {x, y, z};
  double* out[3] = {a1, a2, a3};
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < np1; j++) out[d][j] = 0.0;
  for (int j = 0; j < np1; j++) {
    double dj = (j == 0 || j == P) ? 0.5 : 1.0;
    double w = ((j % 2 == 0) ? 1 : -1) * dj;
    for (int d = 0; d < 3; d++) {
      if (std::fabs(v[d] - T[d][j]) < 1.e-07) {
        flag[d] = j;
      } else {
        double tmp = w / (v[d] - T[d][j]);
        out[d][j] = tmp;
        S[d] += tmp;
      }
    }
  }
  for (int d = 0; d < 3; d++)
    if (flag[d] > -1) {
      S[d] = 1.0;
      for (int j = 0; j < np1; j++) out[d][j] = 0.0;
      out[d][flag[d]] = 1.0;
    }
  *D = 1.0 / (S[0] * S[1] * S[2]);
}

// Moment tables of the nodes flagged in far_flag: barycentric factors of
// every member (host), member-chunk work items, per-item partial and final
// moment buffers (device).  Sets far_idx, nfar_nodes and the plan's
// moment_terms.
static int build_moment_tables(hor_plan* P, const double* sx) {
  hipStream_t s = P->ctx->stream;
  TreePlan& TP = *P->tree;
  const Tree& T = TP.T;
  const bool dl = (P->kind == HOR_DL_ON || P->kind == HOR_DL_OFF);
  const int nn = (int)T.nodes.size(), np1 = TP.np1, stride = 3 * np1 + 1;
  std::vector<int> far_idx(nn + 1, -1), far_nodes;
  for (int q = 0; q < nn; q++)
    if (TP.far_flag[q]) { far_idx[q] = (int)far_nodes.size(); far_nodes.push_back(q); }
  TP.nfar_nodes = (int)far_nodes.size();
  std::vector<int> fac_off(TP.nfar_nodes + 1, 0);
  for (int f = 0; f < TP.nfar_nodes; f++) {
    const TreeNode& nd = T.nodes[far_nodes[f]];
    fac_off[f + 1] = fac_off[f] + (nd.hi - nd.lo + 1) * stride;
  }
  std::vector<double> fac(std::max<size_t>(fac_off[TP.nfar_nodes], 1), 0.0);
  double moment_terms = 0;
  long total_members = 0;
  for (int q : far_nodes) total_members += T.nodes[q].hi - T.nodes[q].lo + 1;
  const int yblocks = (np1 * np1 + 63) / 64;   // k_moments workgroups per item
  const long want_items = std::max(1, kMomWantBlocks / yblocks);
  const int chunk = (int)std::min<long>(kMomChunkMax, std::max<long>(kMomChunkMin,
                                        ((total_members + want_items - 1) / want_items + 63) / 64 * 64));
  std::vector<int> mi_far, mi_m0, mi_m1, m_first(TP.nfar_nodes), m_nit(TP.nfar_nodes);
  for (int f = 0; f < TP.nfar_nodes; f++) {
    const int q = far_nodes[f], lo = T.nodes[q].lo, nm = T.nodes[q].hi - lo + 1;
    m_first[f] = (int)mi_far.size();
    for (int a = 0; a < nm; a += chunk) {
      mi_far.push_back(f);
      mi_m0.push_back(a);
      mi_m1.push_back(std::min(nm, a + chunk));
    }
    m_nit[f] = (int)mi_far.size() - m_first[f];
    double* F = &fac[fac_off[f]];
    const double* cq = &T.cheb[3 * (size_t)np1 * q];
    for (int m = 0; m < nm; m++) {
      const int o = T.perm[lo + m];
      double* fm = F + (size_t)m * stride;
      bary_factors(np1 - 1, cq, cq + np1, cq + 2 * np1, sx[3 * o], sx[3 * o + 1], sx[3 * o + 2], fm, fm + np1,
                   fm + 2 * np1, fm + 3 * np1);
    }
    moment_terms += (double)nm * TP.pflat;
  }
  P->moment_terms = moment_terms;
  TP.n_mom_items = (int)mi_far.size();
  HOR_TRY(TP.far_idx.upload(far_idx, s));
  HOR_TRY(TP.far_node.upload(far_nodes.empty() ? std::vector<int>{0} : far_nodes, s));
  HOR_TRY(TP.mom_item_far.upload(mi_far, s));
  HOR_TRY(TP.mom_item_m0.upload(mi_m0, s));
  HOR_TRY(TP.mom_item_m1.upload(mi_m1, s));
  HOR_TRY(TP.mom_node_first.upload(m_first, s));
  HOR_TRY(TP.mom_node_nitems.upload(m_nit, s));
  HOR_TRY(TP.mom_part.alloc((size_t)std::max(TP.n_mom_items, 1) * TP.ncomp * TP.pflat));
  HOR_TRY(TP.fac_off.upload(fac_off, s));
  HOR_TRY(TP.fac.upload(fac, s));
  HOR_TRY(TP.mom.alloc((size_t)std::max(TP.nfar_nodes, 1) * TP.ncomp * TP.pflat));
  if (P->precision == HOR_MIXED)
    HOR_TRY(TP.mom32.alloc((size_t)std::max(TP.nfar_nodes, 1) * (dl ? kNc32Dl : kNc32Sl) * TP.pflat));
  return HOR_OK;
}

int tree_plan_create(hor_plan* P, const double* sx, const double* sn, const double* sa, const double* tx,
                     const double* tb, const double* tx0, const double* tn0, const double* ctr,
                     const hor_tree_params& tp) {
  hipStream_t s = P->ctx->stream;
  P->tree.reset(new TreePlan());
  TreePlan& TP = *P->tree;
  for (hipEvent_t& e : TP.ev) HOR_HIP_TRY(hipEventCreate(&e));
  Tree& T = TP.T;
  tree_build(T, P->ns, sx, ctr, tp.P, tp.N0, tp.theta);
  TP.np1 = tp.P + 1;
  TP.pflat = TP.np1 * TP.np1 * TP.np1;
  bool dl = (P->kind == HOR_DL_ON || P->kind == HOR_DL_OFF);
  TP.ncomp = dl ? 9 : 6;
  // sources in tree order; targets sorted spatially (the class rule of the
  // treecode's leaf sums: |b|/delta > 8, KITC.cpp:856, 970)
  HOR_TRY(setup_sources(P, T.perm, sx, sn, sa));
  HOR_TRY(setup_targets(P, tx, tb, tx0, tn0, sx));
  const int nt = P->nt, nn = (int)T.nodes.size();
  tree_lists(T, nt, tx, TP.L);
  TP.nslot = TP.L.start[nt];
  // node arrays, leaf subtile boxes (64-source subtiles starting at each node's lo)
  std::vector<int> lo(nn), hi(nn), lso(nn + 1, 0);
  for (int q = 0; q < nn; q++) {
    lo[q] = T.nodes[q].lo;
    hi[q] = T.nodes[q].hi;
    lso[q + 1] = lso[q] + (hi[q] - lo[q] + kWave) / kWave;
  }
  std::vector<double> lbox(6 * (size_t)std::max(lso[nn], 1), 0.0);
  for (int q = 0; q < nn; q++)
    for (int b = 0; b < lso[q + 1] - lso[q]; b++) {
      double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
      for (int p = lo[q] + b * kWave; p <= std::min(hi[q], lo[q] + (b + 1) * kWave - 1); p++) {
        int o = T.perm[p];
        for (int k = 0; k < 3; k++) {
          l[k] = std::min(l[k], sx[3 * o + k]);
          h[k] = std::max(h[k], sx[3 * o + k]);
        }
      }
      for (int k = 0; k < 3; k++) { lbox[6 * (lso[q] + b) + k] = l[k]; lbox[6 * (lso[q] + b) + 3 + k] = h[k]; }
    }
  // transpose the lists into per-node work items, targets in sorted order
  std::vector<std::vector<int>> ntg(nn), nsl(nn), ftg(nn), fsl(nn);
  for (int i = 0; i < nt; i++) {
    int o = P->tperm_h[i];
    for (int q = TP.L.start[o]; q < TP.L.start[o + 1]; q++) {
      int node = TP.L.node[q];
      if (TP.L.far[q]) { ftg[node].push_back(i); fsl[node].push_back(q); }
      else { ntg[node].push_back(i); nsl[node].push_back(q); }
    }
  }
  auto items = [&](std::vector<std::vector<int>>& tg, std::vector<std::vector<int>>& sl, std::vector<int>& inode,
                   std::vector<int>& ioff, std::vector<int>& icnt, std::vector<int>& atg, std::vector<int>& asl) {
    for (int q = 0; q < nn; q++) {
      int m = (int)tg[q].size();
      for (int a = 0; a < m; a += kBlock) {
        inode.push_back(q);
        ioff.push_back((int)atg.size() + a);
        icnt.push_back(std::min(kBlock, m - a));
      }
      atg.insert(atg.end(), tg[q].begin(), tg[q].end());
      asl.insert(asl.end(), sl[q].begin(), sl[q].end());
    }
  };
  std::vector<int> ni_node, ni_off, ni_cnt, n_tg, n_sl, fi_node, fi_off, fi_cnt, f_tg, f_sl;
  items(ntg, nsl, ni_node, ni_off, ni_cnt, n_tg, n_sl);
  items(ftg, fsl, fi_node, fi_off, fi_cnt, f_tg, f_sl);
  // With few near items (a target set near only a few leaves, as the other
  // surface's points are) the leaves' sources are split into chunks of whole
  // 64-source subtiles so that the near sum still has about kNearWantItems
  // workgroups; each chunk of a (target, leaf) entry gets a slot of its own,
  // consecutive.
  const int split = std::max(1, std::min(kNearMaxSplit, (kNearWantItems + std::max(1, (int)ni_node.size()) - 1) /
                                                            std::max(1, (int)ni_node.size())));
  std::vector<int> nchunk(nn, 1), csub(nn, 1);
  for (int q = 0; q < nn; q++) {
    const int nsub = lso[q + 1] - lso[q];
    const int c = std::max(1, std::min(split, nsub));
    csub[q] = (nsub + c - 1) / c;               // subtiles per chunk
    nchunk[q] = (nsub + csub[q] - 1) / csub[q];
  }
  // slots of each list entry, targets in sorted order (the finalize kernel's
  // wavefronts), a target's entries contiguous in visiting order
  std::vector<int> new_start(TP.L.node.size(), 0), slot_start(nt + 1, 0);
  for (int i = 0, next = 0; i < nt; i++) {
    const int o = P->tperm_h[i];
    slot_start[i] = next;
    for (int e = TP.L.start[o]; e < TP.L.start[o + 1]; e++) {
      new_start[e] = next;
      next += TP.L.far[e] ? 1 : nchunk[TP.L.node[e]];
    }
    slot_start[i + 1] = next;
  }
  for (auto& v : n_sl) v = new_start[v];
  for (auto& v : f_sl) v = new_start[v];
  TP.nslot = slot_start[nt];
  std::vector<int> ci_node, ci_off, ci_cnt, ci_src;
  for (size_t it = 0; it < ni_node.size(); it++) {
    const int q = ni_node[it];
    for (int c = 0; c < nchunk[q]; c++) {
      ci_node.push_back(q);
      ci_off.push_back(ni_off[it]);
      ci_cnt.push_back(ni_cnt[it]);
      const int j0 = lo[q] + c * csub[q] * kWave, j1 = std::min(hi[q] + 1, j0 + csub[q] * kWave);
      ci_src.insert(ci_src.end(), {j0, j1, lso[q] + c * csub[q], c});
    }
  }
  ni_node.swap(ci_node);
  ni_off.swap(ci_off);
  ni_cnt.swap(ci_cnt);
  TP.n_near_items = (int)ni_node.size();
  TP.n_far_items = (int)fi_node.size();
  double near_pairs = 0;
  for (int q = 0; q < nn; q++) near_pairs += (double)ntg[q].size() * (hi[q] - lo[q] + 1);
  P->near_pairs = near_pairs;
  P->proxy_pairs = (double)f_tg.size() * TP.pflat;
  TP.far_flag.assign(nn, 0);
  for (int q = 0; q < nn; q++) TP.far_flag[q] = !ftg[q].empty();
  P->direct_pairs = (double)P->ns * P->nt;
  // uploads
  HOR_TRY(TP.node_lo.upload(lo, s));
  HOR_TRY(TP.node_hi.upload(hi, s));
  HOR_TRY(TP.leaf_sub_off.upload(lso, s));
  HOR_TRY(TP.leaf_box.upload(lbox, s));
  {
    // each slot's leaf subtile: nodes are numbered parents first, so the last
    // node holding a slot is its leaf
    std::vector<int> ssub(P->nsp, 0);
    for (int q = 0; q < nn; q++)
      for (int j = lo[q]; j <= hi[q]; j++) ssub[j] = lso[q] + (j - lo[q]) / kWave;
    HOR_TRY(TP.slot_sub.upload(ssub, s));
  }
  HOR_TRY(TP.cheb.upload(T.cheb, s));
  {
    // node centres (midpoints of the Chebyshev-Lobatto intervals) and the
    // node-centred FP32 points of the mixed-precision far field
    const int np1 = TP.np1;
    std::vector<double> nctr(3 * (size_t)nn);
    std::vector<float> c32(T.cheb.size());
    for (int q = 0; q < nn; q++)
      for (int d = 0; d < 3; d++) {
        const double* t = &T.cheb[(3 * (size_t)q + d) * np1];
        const double cm = 0.5 * (t[0] + t[np1 - 1]);
        nctr[3 * (size_t)q + d] = cm;
        for (int j = 0; j < np1; j++) c32[(3 * (size_t)q + d) * np1 + j] = (float)(t[j] - cm);
      }
    HOR_TRY(TP.node_ctr.upload(nctr, s));
    HOR_TRY(TP.cheb32.upload(c32, s));
  }
  HOR_TRY(TP.near_item_node.upload(ni_node, s));
  HOR_TRY(TP.near_item_off.upload(ni_off, s));
  HOR_TRY(TP.near_item_cnt.upload(ni_cnt, s));
  HOR_TRY(TP.near_item_src.upload(ci_src.empty() ? std::vector<int>{0, 0, 0, 0} : ci_src, s));
  HOR_TRY(TP.near_tgt.upload(n_tg, s));
  HOR_TRY(TP.near_slot.upload(n_sl, s));
  HOR_TRY(TP.far_item_node.upload(fi_node, s));
  HOR_TRY(TP.far_item_off.upload(fi_off, s));
  HOR_TRY(TP.far_item_cnt.upload(fi_cnt, s));
  HOR_TRY(TP.far_tgt.upload(f_tg, s));
  HOR_TRY(TP.far_slot.upload(f_sl, s));
  HOR_TRY(TP.mq.alloc((size_t)P->nsp * 16));
  HOR_TRY(TP.slot_start.upload(slot_start, s));
  HOR_TRY(TP.slots.alloc(3 * (size_t)std::max(TP.nslot, 1)));
  HOR_TRY(build_moment_tables(P, sx));
  // xi of the treecode off-surface double layer (set by setup_targets for
  // method == HOR_TREE) uses the given g0 of every target.
  return HOR_OK;
}

// In the two-surface matvec the on-surface plan of a surface and its plan at
// the other surface's points sum over the same source tree with the same
// density, one after the other: the first (owner) computes the moments of
// both plans' far nodes and the second (user) reads them.  The trees must be
// identical (same sources, tree parameters and centre).
int tree_plan_share_moments(hor_plan* owner, hor_plan* user, const double* sx) {
  if (!owner || !user || !owner->tree || !user->tree)
    return set_error(HOR_E_ARG, "moment sharing needs two treecode plans");
  auto is_dl = [](int k) { return k == HOR_DL_ON || k == HOR_DL_OFF; };
  TreePlan& O = *owner->tree;
  TreePlan& U = *user->tree;
  if (is_dl(owner->kind) != is_dl(user->kind) || owner->precision != user->precision ||
      owner->ns != user->ns || O.np1 != U.np1 || O.T.nodes.size() != U.T.nodes.size() || O.T.perm != U.T.perm ||
      O.T.cheb != U.T.cheb || U.mom_owner || O.mom_owner)
    return set_error(HOR_E_ARG, "moment sharing: the plans' source trees differ");
  for (size_t q = 0; q < O.T.nodes.size(); q++)
    if (O.T.nodes[q].lo != U.T.nodes[q].lo || O.T.nodes[q].hi != U.T.nodes[q].hi)
      return set_error(HOR_E_ARG, "moment sharing: the plans' source trees differ");
  for (size_t q = 0; q < O.far_flag.size(); q++) O.far_flag[q] = O.far_flag[q] || U.far_flag[q];
  HOR_TRY(build_moment_tables(owner, sx));
  U.mom_owner = &O;
  user->moment_terms = 0;
  U.mom.release();
  U.mom32.release();
  U.mom_part.release();
  U.fac.release();
  return HOR_OK;
}

template <int KIND>
static void launch_leaf(const LeafArgs& A, bool mixed, bool nosub, int nitems, hipStream_t s) {
  const size_t lds = sizeof(double) * table_len(KIND);
  constexpr bool kDl = KIND == HOR_DL_ON || KIND == HOR_DL_OFF;
  if (kDl && nosub) {
    if (mixed) hipLaunchKernelGGL((k_leaf<KIND, true, kDl>), dim3(nitems), dim3(kBlock), lds, s, A);
    else hipLaunchKernelGGL((k_leaf<KIND, false, kDl>), dim3(nitems), dim3(kBlock), lds, s, A);
  } else {
    if (mixed) hipLaunchKernelGGL((k_leaf<KIND, true, false>), dim3(nitems), dim3(kBlock), lds, s, A);
    else hipLaunchKernelGGL((k_leaf<KIND, false, false>), dim3(nitems), dim3(kBlock), lds, s, A);
  }
}

// One evaluation of a treecode plan.  column < 0: the plan's operator on
// dens (d0 the targets' given densities), written as base/coef/accumulate
// say.  column l in 0..2: column l of the double layer's subtraction
// coefficients (TreePlan::csub), FP64, into out (3 nt, the caller's order):
// zero density, g0 = -e_l, no jump term.
static int tree_eval(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev, const double* base,
                     double coef, int accumulate, int column) {
  hor_ctx* c = P->ctx;
  hipStream_t s = c->stream;
  TreePlan& TP = *P->tree;
  const bool coefs = column >= 0;
  const bool mixed = P->precision == HOR_MIXED && !coefs;
  const bool dl = (P->kind == HOR_DL_ON || P->kind == HOR_DL_OFF);
  const bool nosub = dl && TP.has_csub && !coefs;
  // density in tree order; per-target subtraction data (leaf and proxy variants)
  HOR_TRY(launch_gather3(dens_dev, P->sperm.p, P->ns, P->nsp, P->sx.p, P->sn.p, P->sa.p, P->sd.p, P->rec.p,
                         mixed ? P->rec32.p : nullptr, TP.leaf_box.p,
                         P->kind == HOR_SL_ON || P->kind == HOR_SL_OFF, s, TP.slot_sub.p));
  if (coefs) {
    hipLaunchKernelGGL(k_csub_targets, dim3((P->ntp + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P->ntp, column,
                       P->kind == HOR_DL_OFF ? 1 : 0, P->tnear.p, P->tsub.p, P->tsub2.p);
    HOR_HIP_TRY(hipGetLastError());
  } else {
    HOR_TRY(prepare_targets(P, dens_dev, d0_dev));
    if (P->kind == HOR_SL_OFF)
      HOR_TRY(launch_target_f0n0_all(P->nt, P->ntp, P->tperm.p, d0_dev, P->n0_orig.p, P->tsub2.p, s));
    else if (P->kind == HOR_DL_OFF)
      HOR_TRY(launch_target_sub(4, P->nt, P->ntp, P->tperm.p, P->tnear.p, d0_dev, nullptr, nullptr, nullptr,
                                P->tsub2.p, s));
    else
      HOR_HIP_TRY(hipMemcpyAsync(P->tsub2.p, P->tsub.p, sizeof(double) * 3 * P->ntp, hipMemcpyDeviceToDevice, s));
  }
  HOR_HIP_TRY(hipEventRecord(c->ev[2], s));
  HOR_HIP_TRY(hipEventRecord(TP.ev[0], s));
  // 1. moments (unless another plan computes them)
  if (TP.nfar_nodes > 0 && !TP.mom_owner) {
    hipLaunchKernelGGL(k_moment_rhs, dim3((P->ns + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P->ns, P->nsp,
                       dl ? 1 : 0, P->sn.p, P->sa.p, P->sd.p, TP.mq.p);
    MomArgs M;
    M.np1 = TP.np1; M.pflat = TP.pflat; M.ncomp = TP.ncomp;
    M.item_far = TP.mom_item_far.p; M.item_m0 = TP.mom_item_m0.p; M.item_m1 = TP.mom_item_m1.p;
    M.far_node = TP.far_node.p; M.node_lo = TP.node_lo.p; M.fac_off = TP.fac_off.p;
    M.fac = TP.fac.p; M.mq = TP.mq.p; M.part = TP.mom_part.p;
    const int rows = TP.np1 * TP.np1;   // 16 rows per wave, 4 waves per workgroup
    const dim3 mg(TP.n_mom_items, (rows + 63) / 64);
    switch ((TP.np1 * TP.ncomp + 15) / 16) {   // column tiles, 1..kMomTiles for P <= 12
      case 1: hipLaunchKernelGGL(k_moments<1>, mg, dim3(kBlock), 0, s, M); break;
      case 2: hipLaunchKernelGGL(k_moments<2>, mg, dim3(kBlock), 0, s, M); break;
      case 3: hipLaunchKernelGGL(k_moments<3>, mg, dim3(kBlock), 0, s, M); break;
      case 4: hipLaunchKernelGGL(k_moments<4>, mg, dim3(kBlock), 0, s, M); break;
      case 5: hipLaunchKernelGGL(k_moments<5>, mg, dim3(kBlock), 0, s, M); break;
      case 6: hipLaunchKernelGGL(k_moments<6>, mg, dim3(kBlock), 0, s, M); break;
      case 7: hipLaunchKernelGGL(k_moments<7>, mg, dim3(kBlock), 0, s, M); break;
      default: hipLaunchKernelGGL(k_moments<kMomTiles>, mg, dim3(kBlock), 0, s, M); break;
    }
    const int rblocks = (TP.ncomp * TP.pflat + kMomReduceBlock - 1) / kMomReduceBlock;
    hipLaunchKernelGGL(k_moments_reduce, dim3(TP.nfar_nodes, rblocks), dim3(kMomReduceBlock), 0, s, TP.ncomp,
                       TP.pflat,
                       TP.mom_node_first.p, TP.mom_node_nitems.p, TP.mom_part.p, TP.mom.p,
                       mixed ? TP.mom32.p : nullptr, dl ? kNc32Dl : kNc32Sl);
    HOR_HIP_TRY(hipGetLastError());
  }
  HOR_HIP_TRY(hipEventRecord(TP.ev[1], s));
  // 2. near leaves
  if (TP.n_near_items > 0) {
    LeafArgs A;
    A.D = direct_args(P);
    A.D.sub_box = TP.leaf_box.p;
    A.item_off = TP.near_item_off.p; A.item_cnt = TP.near_item_cnt.p;
    A.tgt = TP.near_tgt.p; A.slot = TP.near_slot.p;
    A.item_src = TP.near_item_src.p;
    A.slots = TP.slots.p; A.nslot = TP.nslot;
    switch (P->kind) {
      case HOR_SL_ON: launch_leaf<HOR_SL_ON>(A, mixed, false, TP.n_near_items, s); break;
      case HOR_SL_OFF: launch_leaf<HOR_SL_OFF>(A, mixed, false, TP.n_near_items, s); break;
      case HOR_DL_ON: launch_leaf<HOR_DL_ON>(A, mixed, nosub, TP.n_near_items, s); break;
      default: launch_leaf<HOR_DL_OFF>(A, mixed, nosub, TP.n_near_items, s); break;
    }
    HOR_HIP_TRY(hipGetLastError());
  }
  HOR_HIP_TRY(hipEventRecord(TP.ev[2], s));
  // 3. far field
  if (TP.n_far_items > 0) {
    FarArgs F;
    F.np1 = TP.np1; F.pflat = TP.pflat; F.ntp = P->ntp; F.nslot = TP.nslot;
    F.item_node = TP.far_item_node.p; F.item_off = TP.far_item_off.p; F.item_cnt = TP.far_item_cnt.p;
    const TreePlan& MO = TP.mom_owner ? *TP.mom_owner : TP;
    F.tgt = TP.far_tgt.p; F.slot = TP.far_slot.p; F.far_idx = MO.far_idx.p;
    F.cheb = TP.cheb.p; F.mom = MO.mom.p; F.tx = P->tx.p; F.tsub2 = P->tsub2.p; F.slots = TP.slots.p;
    F.node_ctr = TP.node_ctr.p; F.cheb32 = TP.cheb32.p; F.mom32 = MO.mom32.p;
    dim3 grid(TP.n_far_items), b64(kBlock / kFarT64), b32(kBlock / kFarT32);
    if (mixed && !dl && TP.far_kernel == HOR_FAR_MFMA) {
      const int pflat4 = (TP.pflat + 3) & ~3;
      const size_t need = (size_t)std::max(MO.nfar_nodes, 1) * pflat4 * kFarMfmaCols;
      if (TP.bmat.n < need) HOR_TRY(TP.bmat.alloc(need));
      if (MO.nfar_nodes > 0)
        hipLaunchKernelGGL(k_far_sl_bmat, dim3((pflat4 + kBlock - 1) / kBlock, MO.nfar_nodes), dim3(kBlock), 0, s,
                           MO.nfar_nodes, TP.np1, TP.pflat, pflat4, MO.far_node.p, TP.cheb32.p, MO.mom32.p,
                           TP.bmat.p);
      FarMfmaArgs G;
      G.pflat4 = pflat4; G.ntp = P->ntp; G.nslot = TP.nslot;
      G.item_node = F.item_node; G.item_off = F.item_off; G.item_cnt = F.item_cnt;
      G.tgt = F.tgt; G.slot = F.slot; G.far_idx = F.far_idx;
      G.tx = F.tx; G.tsub2 = F.tsub2; G.node_ctr = F.node_ctr; G.B = TP.bmat.p; G.slots = F.slots;
      hipLaunchKernelGGL(k_far_sl_mfma, grid, dim3(kBlock), 0, s, G);
    } else if (dl && nosub) {
      if (mixed) hipLaunchKernelGGL((k_far<true, true, kFarT32, true>), grid, b32, 0, s, F);
      else hipLaunchKernelGGL((k_far<true, false, kFarT64, true>), grid, b64, 0, s, F);
    } else if (dl) {
      if (mixed) hipLaunchKernelGGL((k_far<true, true, kFarT32, false>), grid, b32, 0, s, F);
      else hipLaunchKernelGGL((k_far<true, false, kFarT64, false>), grid, b64, 0, s, F);
    } else {
      if (mixed) hipLaunchKernelGGL((k_far<false, true, kFarT32, false>), grid, b32, 0, s, F);
      else hipLaunchKernelGGL((k_far<false, false, kFarT64, false>), grid, b64, 0, s, F);
    }
    HOR_HIP_TRY(hipGetLastError());
  }
  HOR_HIP_TRY(hipEventRecord(c->ev[3], s));
  HOR_HIP_TRY(hipEventRecord(TP.ev[3], s));
  TP.evaluated = true;
  // 4. per-target reduction
  hipLaunchKernelGGL(k_tree_finalize, dim3((P->nt + kBlock / kFinLanes - 1) / (kBlock / kFinLanes)), dim3(kBlock), 0, s,
                     P->nt, P->ntp,
                     TP.nslot, P->tperm.p, TP.slot_start.p, TP.slots.p,
                     dl ? -6.0 / (8.0 * kPi) : 1.0 / (8.0 * kPi),   // KITC.cpp:1395, 1612
                     dl && !coefs ? P->xi.p : nullptr, dl ? P->tsub2.p : nullptr, nosub ? TP.csub.p : nullptr,
                     out_dev, base, coef, accumulate);
  HOR_HIP_TRY(hipGetLastError());
  return HOR_OK;
}

static int tree_plan_csub(hor_plan* P);

int tree_plan_eval(hor_plan* P, const double* dens_dev, const double* d0_dev, double* out_dev, const double* base,
                   double coef, int accumulate) {
  TreePlan& TP = *P->tree;
  const bool dl = P->kind == HOR_DL_ON || P->kind == HOR_DL_OFF;
  if (dl && !TP.has_csub && !TP.mom_owner && ++TP.n_evals == 2) HOR_TRY(tree_plan_csub(P));
  return tree_eval(P, dens_dev, d0_dev, out_dev, base, coef, accumulate, -1);
}

int tree_plan_prepare_subtraction(hor_plan* P) {
  if (!P || !P->tree) return set_error(HOR_E_ARG, "subtraction coefficients: not a treecode plan");
  if (P->kind != HOR_DL_ON && P->kind != HOR_DL_OFF) return HOR_OK;
  if (P->tree->has_csub) return HOR_OK;
  if (P->tree->mom_owner)
    return set_error(HOR_E_STATE, "subtraction coefficients: the plan reads another plan's moments");
  return tree_plan_csub(P);
}

// The double layer's subtraction coefficients (TreePlan::csub) of a treecode
// plan that computes its own moments: three FP64 evaluations.
static int tree_plan_csub(hor_plan* P) {
  TreePlan& TP = *P->tree;
  if (P->nt == 0) {
    TP.has_csub = true;
    return HOR_OK;
  }
  hipStream_t s = P->ctx->stream;
  const int nt = P->nt;
  DBuf<double> zero, col;
  HOR_TRY(zero.alloc(3 * (size_t)std::max(P->ns, 1)));
  HOR_TRY(col.alloc(3 * (size_t)nt));
  HOR_TRY(TP.csub.alloc(9 * (size_t)P->ntp));
  HOR_HIP_TRY(hipMemsetAsync(zero.p, 0, sizeof(double) * zero.n, s));
  HOR_HIP_TRY(hipMemsetAsync(TP.csub.p, 0, sizeof(double) * TP.csub.n, s));
  for (int l = 0; l < 3; l++) {
    HOR_TRY(tree_eval(P, zero.p, nullptr, col.p, nullptr, 1.0, 0, l));
    hipLaunchKernelGGL(k_csub_store, dim3((nt + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nt, P->ntp, l,
                       P->tperm.p, col.p, TP.csub.p);
    HOR_HIP_TRY(hipGetLastError());
  }
  // the buffers above are freed on return: the stream must be done with them
  HOR_HIP_TRY(hipStreamSynchronize(s));
  TP.has_csub = true;
  TP.evaluated = false;
  return HOR_OK;
}

int tree_plan_set_far_kernel(hor_plan* P, int kernel) {
  if (!P || !P->tree) return set_error(HOR_E_ARG, "far-field kernel: not a treecode plan");
  if (kernel != HOR_FAR_VALU && kernel != HOR_FAR_MFMA) return set_error(HOR_E_ARG, "far-field kernel: bad value");
  if (kernel == HOR_FAR_MFMA &&
      (P->precision != HOR_MIXED || (P->kind != HOR_SL_ON && P->kind != HOR_SL_OFF)))
    return set_error(HOR_E_ARG, "the matrix-core far field is for mixed-precision single-layer treecode plans");
  P->tree->far_kernel = kernel;
  return HOR_OK;
}

// Kernel time of each phase of the plan's last evaluation (HIP events on the
// context's stream; call after the stream has synchronized): moments (moment
// right-hand side, GEMM and reduction), near sums (k_leaf), far field (k_far).
int tree_plan_phase_ms(hor_plan* P, double ms[3]) {
  if (!P->tree) return set_error(HOR_E_ARG, "not a treecode plan");
  const TreePlan& TP = *P->tree;
  if (!TP.evaluated) return set_error(HOR_E_STATE, "phase timing before the plan's first evaluation");
  for (int k = 0; k < 3; k++) {
    float t = 0;
    HOR_HIP_TRY(hipEventElapsedTime(&t, TP.ev[k], TP.ev[k + 1]));
    ms[k] = t;
  }
  return HOR_OK;
}

}  // namespace hor

hor_plan::hor_plan() = default;
hor_plan::~hor_plan() = default;
