// A plain device copy of known size (512 MiB read + 512 MiB written per
// launch).  Profiled under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, it gives
// the factor between what those counters report and the bytes actually moved,
// which scales the counters read for the sum kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_calib(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main() {
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
