// Calibration kernels of known work for the rocprofv3 counters.
//   (no argument) a plain device copy of known size (512 MiB read + 512 MiB
//     written per launch): the factor between FETCH_SIZE / WRITE_SIZE and the
//     bytes actually moved, which scales the counters read for the sum kernels;
//   "valu": FP64 and FP32 fused multiply-add chains of known length
//     (16384 waves x 4096 FMAs per launch and precision): the factor between
//     SQ_INSTS_VALU_FMA_F64 / _F32 and the wave-level FMA instructions issued.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

constexpr int kIters = 1024;  // x 4 independent chains = 4096 FMAs per thread

__device__ inline double fm(double x, double a, double b) { return __builtin_fma(x, a, b); }
__device__ inline float fm(float x, float a, float b) { return __builtin_fmaf(x, a, b); }

template <class T>
__global__ __launch_bounds__(256) void fma_calib(T a, T b, T* out) {
  T x0 = (T)threadIdx.x, x1 = x0 + (T)1, x2 = x0 + (T)2, x3 = x0 + (T)3;
  for (int k = 0; k < kIters; k++) {
    x0 = fm(x0, a, b);
    x1 = fm(x1, a, b);
    x2 = fm(x2, a, b);
    x3 = fm(x3, a, b);
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3;
}

static int run_valu() {
  const size_t n = (size_t)4096 * 256;
  double* d = nullptr;
  float* f = nullptr;
  if (hipMalloc(&d, n * sizeof(double)) != hipSuccess || hipMalloc(&f, n * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "fma_calib: hipMalloc failed\n");
    return 1;
  }
  for (int r = 0; r < 3; r++) {
    hipLaunchKernelGGL(fma_calib<double>, dim3(4096), dim3(256), 0, 0, 0.999, 1e-3, d);
    hipLaunchKernelGGL(fma_calib<float>, dim3(4096), dim3(256), 0, 0, 0.999f, 1e-3f, f);
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    std::fprintf(stderr, "fma_calib: kernel failed\n");
    return 1;
  }
  std::printf("fma_calib: 16384 waves x %d FMAs per launch (FP64 and FP32), 3 launches each\n", 4 * kIters);
  (void)hipFree(d);
  (void)hipFree(f);
  return 0;
}

__global__ void copy_calib(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "valu") == 0) return run_valu();
  const size_t n = (size_t)64 << 20;  // 64 Mi doubles = 512 MiB each way
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess) {
    std::fprintf(stderr, "copy_calib: hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, n * 8) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 1;
  for (int r = 0; r < 3; r++) hipLaunchKernelGGL(copy_calib, dim3(4096), dim3(256), 0, 0, a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::fprintf(stderr, "copy_calib: kernel failed\n");
    return 1;
  }
  std::printf("copy_calib: %zu bytes read and %zu bytes written per launch, 3 launches\n", n * 8, n * 8);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
