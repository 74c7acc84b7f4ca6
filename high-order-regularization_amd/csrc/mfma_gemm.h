// BF16 matrix-core GEMM for the point-cloud network (gfx950).
//
//   C[M x N] (+)= sum_k A(i, k) B(k, j),
//   A(i, k) = TA ? A[k*lda + i] : A[i*lda + k],  B(k, j) = TB ? B[j*ldb + k] : B[k*ldb + j]
//
// Operands are FP32 in memory and rounded to BF16 (round to nearest even)
// when they are loaded; products accumulate in FP32 in the matrix core
// (v_mfma_f32_32x32x8_bf16).  One wave computes one 32x32 tile of C; the K
// range of blockIdx.z is [z*kchunk, min(K, (z+1)*kchunk)) (kchunk a multiple
// of 8) and its partial goes to C + z*split_stride, so that a split-K product
// is summed afterwards in a fixed order.
//
// Operand layout of the 32x32x8 BF16 instruction (wave64): lane l holds
// A[l%32][4*(l/32) + v] and B[4*(l/32) + v][l%32] for v = 0..3; accumulator
// register r of lane l is C[8*(r/4) + 4*(l/32) + r%4][l%32].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hor {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(64) void k_gemm_bf16(int M, int N, int K, int kchunk, const float* __restrict__ A,
                                                  int lda, const float* __restrict__ B, int ldb, float* C, int ldc,
                                                  size_t split_stride, int accumulate) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int ra = i0 + (lane & 31);
  const int cb = j0 + (lane & 31);
  const int kq = 4 * (lane >> 5);
  f32x16_t acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  for (int k = kb; k < ke; k += 8) {
    bf16x4_t a, b;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int kk = k + kq + v;
      float av = 0.0f, bv = 0.0f;
      if (ra < M && kk < ke) av = TA ? A[(size_t)kk * lda + ra] : A[(size_t)ra * lda + kk];
      if (cb < N && kk < ke) bv = TB ? B[(size_t)cb * ldb + kk] : B[(size_t)kk * ldb + cb];
      a[v] = f32_to_bf16(av);
      b[v] = f32_to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  float* Cz = C + (size_t)blockIdx.z * split_stride;
  const int j = j0 + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i < M && j < N) {
      float* c = Cz + (size_t)i * ldc + j;
      if (accumulate) *c += acc[r];
      else *c = acc[r];
    }
  }
}

}  // namespace hor
