// Plain device-to-device copy of a known size, used to calibrate the HBM
// byte counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports for the
// relaxation kernels, and to measure the achievable copy bandwidth.
//   copy_calib [MiB] [repeats]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

// One 16-byte element per thread over the whole array (the fastest plain
// copy form measured on MI355X: about 6.3 TB/s read + write for 1 GiB).
__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int main(int argc, char** argv) {
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const long long bytes = mib << 20;
  const long long n = bytes / (long long)sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const unsigned blocks = (unsigned)((n + 255) / 256);
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
