// Plain streaming kernels of known size: they calibrate the HBM byte
// counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports for the
// relaxation kernels, and measure the bandwidth a simple access mix reaches.
//   copy_calib [MiB] [repeats] [copy|read|write|triad]
// copy: 1 read + 1 write; read: reads only; write: writes only; triad:
// 3 reads + 1 write per element (the read/write mix of the P kernel).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

// One 16-byte element per thread over the whole array (the fastest plain
// copy form measured on MI355X: about 6.3 TB/s read + write for 1 GiB).
__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void read_kernel(const double2* __restrict__ src,
                                                   double* __restrict__ sink, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 v = src[i];
    if (v.x == 1234.5 && v.y == -1234.5) sink[0] = v.x;  // zero data: never taken
  }
}

__global__ __launch_bounds__(256) void write_kernel(double2* __restrict__ dst, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = double2{1.0, 2.0};
}

__global__ __launch_bounds__(256) void triad_kernel(const double2* __restrict__ a,
                                                    const double2* __restrict__ b,
                                                    const double2* __restrict__ c,
                                                    double2* __restrict__ d, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 x = a[i], y = b[i], z = c[i];
    d[i] = double2{x.x + 0.5 * y.x - z.x, x.y + 0.5 * y.y - z.y};
  }
}

int main(int argc, char** argv) {
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const char* mode = argc > 3 ? argv[3] : "copy";
  const long long bytes = mib << 20;
  const long long n = bytes / (long long)sizeof(double2);
  const bool triad = !strcmp(mode, "triad");
  double2 *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&d, bytes) != hipSuccess) return 1;
  if (triad && (hipMalloc(&b, bytes) != hipSuccess || hipMalloc(&c, bytes) != hipSuccess)) return 1;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(d, 0, bytes) != hipSuccess) return 1;
  if (triad && (hipMemset(b, 0, bytes) != hipSuccess || hipMemset(c, 0, bytes) != hipSuccess))
    return 1;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  double moved = 0;
  auto launch = [&]() {
    if (!strcmp(mode, "read")) {
      hipLaunchKernelGGL(read_kernel, dim3(blocks), dim3(256), 0, 0, a, (double*)d, n);
      moved = (double)bytes;
    } else if (!strcmp(mode, "write")) {
      hipLaunchKernelGGL(write_kernel, dim3(blocks), dim3(256), 0, 0, d, n);
      moved = (double)bytes;
    } else if (triad) {
      hipLaunchKernelGGL(triad_kernel, dim3(blocks), dim3(256), 0, 0, a, b, c, d, n);
      moved = 4.0 * bytes;
    } else {
      hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, d, n);
      moved = 2.0 * bytes;
    }
  };
  launch();
  hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) launch();
  hipEventRecord(e1, 0);
  if (hipEventSynchronize(e1) != hipSuccess) return 1;
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  const double per = ms / reps;
  printf("%s %lld MiB: %.4f ms per pass, %.1f GB/s\n", mode, mib, per, moved / (per * 1e-3) / 1e9);
  hipFree(a);
  hipFree(d);
  if (b) hipFree(b);
  if (c) hipFree(c);
  return 0;
}
