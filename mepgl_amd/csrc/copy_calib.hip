// Plain device-to-device copy of a known size, used to calibrate the HBM
// byte counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports for the
// relaxation kernels, and to measure the achievable copy bandwidth.
//   copy_calib [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src,
                                                   double4* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

int main(int argc, char** argv) {
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const long long bytes = mib << 20;
  const long long n = bytes / (long long)sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const unsigned blocks = 256 * 32;
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  const double per = ms / reps;
  printf("copy %lld MiB: %.4f ms per copy, %.1f GB/s (read + write)\n", mib, per,
         2.0 * bytes / (per * 1e-3) / 1e9);
  hipFree(a);
  hipFree(b);
  return 0;
}
