// Plain streaming kernels of known size: they calibrate the HBM byte
// counters (FETCH_SIZE, WRITE_SIZE) that rocprofv3 reports for the
// relaxation kernels, and measure the bandwidth a simple access mix reaches
// at a given access width per lane.
//   copy_calib [MiB] [repeats] [copy|read|write|triad] [bytes per lane]
// copy: 1 read + 1 write; read: reads only; write: writes only; triad:
// 3 reads + 1 write per element (the read/write mix of the P kernel), one
// element per thread; stream: the same mix from a grid of [waves per SIMD]
// wavefronts per SIMD (default 4, as the mixed stencil kernels run) that
// loops over the arrays, each wavefront issuing the loads of 4 elements of
// every array before their stores (the stencil kernels' way of keeping
// loads in flight).
//   copy_calib [MiB] [repeats] [mode] [bytes per lane] [waves per SIMD]
// Bytes per lane: 4 (one dword: the mixed-precision stencil kernels' load
// width with one column per lane), 8 (dwordx2: FP64 with one column per
// lane, or mixed with two), 16 (dwordx4, the default).
//   copy_calib [images] [repeats] step [4|8] [segment rows]
// step: the access pattern of one NLCG step of the relaxation kernels on
// `images` one-component images of 256 x 256 sites (4 planes interleaved by
// row), without their arithmetic: a G-like launch reads x, d and g_old and
// writes x' and g, a P-like launch reads g, d and x' and writes d', in the
// kernels' alternating order over the same five field sets.  Each
// 4-wavefront workgroup marches `segment rows` rows of four adjacent 64-column
// strips (one wavefront per strip, one element per lane and plane), as the
// mixed-precision kernels do; edge columns are not loaded, as there.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

// One element of W bytes per thread over the whole array.
template <int W>
struct Elem;
template <>
struct Elem<4> {
  using T = float;
  __device__ static T val(float a) { return a; }
  __device__ static T mix(T x, T y, T z) { return x + 0.5f * y - z; }
  __device__ static bool hit(T v) { return v == 1234.5f; }
};
template <>
struct Elem<8> {
  using T = double;
  __device__ static T val(float a) { return a; }
  __device__ static T mix(T x, T y, T z) { return x + 0.5 * y - z; }
  __device__ static bool hit(T v) { return v == 1234.5; }
};
template <>
struct Elem<16> {
  using T = double2;
  __device__ static T val(float a) { return double2{a, 2.0 * a}; }
  __device__ static T mix(T x, T y, T z) {
    return double2{x.x + 0.5 * y.x - z.x, x.y + 0.5 * y.y - z.y};
  }
  __device__ static bool hit(T v) { return v.x == 1234.5 && v.y == -1234.5; }
};

template <int W>
__global__ __launch_bounds__(256) void copy_kernel(const typename Elem<W>::T* __restrict__ src,
                                                   typename Elem<W>::T* __restrict__ dst,
                                                   long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template <int W>
__global__ __launch_bounds__(256) void read_kernel(const typename Elem<W>::T* __restrict__ src,
                                                   double* __restrict__ sink, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) {
    const typename Elem<W>::T v = src[i];
    if (Elem<W>::hit(v)) sink[0] = 1.0;  // zero data: never taken
  }
}

template <int W>
__global__ __launch_bounds__(256) void write_kernel(typename Elem<W>::T* __restrict__ dst,
                                                    long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = Elem<W>::val(1.0f);
}

template <int W>
__global__ __launch_bounds__(256) void triad_kernel(const typename Elem<W>::T* __restrict__ a,
                                                    const typename Elem<W>::T* __restrict__ b,
                                                    const typename Elem<W>::T* __restrict__ c,
                                                    typename Elem<W>::T* __restrict__ d,
                                                    long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) d[i] = Elem<W>::mix(a[i], b[i], c[i]);
}

template <int W>
__global__ __launch_bounds__(256) void stream_kernel(const typename Elem<W>::T* __restrict__ a,
                                                     const typename Elem<W>::T* __restrict__ b,
                                                     const typename Elem<W>::T* __restrict__ c,
                                                     typename Elem<W>::T* __restrict__ d,
                                                     long long n) {
  using T = typename Elem<W>::T;
  constexpr int U = 4;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x; i0 < n; i0 += U * stride) {
    T x[U], y[U], z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        x[u] = a[i];
        y[u] = b[i];
        z[u] = c[i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) d[i] = Elem<W>::mix(x[u], y[u], z[u]);
    }
  }
}

// One row segment of one workgroup: per row, every lane loads its column of
// the 4 planes of R field sets and stores the 4 planes of NW field sets.
template <class T, int R, int NW>
struct StepBufs {
  const T* in[R];
  T* out[NW];
};

template <class T, int R, int NW>
__global__ __launch_bounds__(256) void step_kernel(StepBufs<T, R, NW> p, int h) {
  constexpr long long kRow = 4 * 256;  // elements per row of one field set
  for (int r = 0; r < h; ++r) {
    const long long base = ((long long)blockIdx.x * h + r) * kRow + threadIdx.x;
    T v[R][4];
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) v[k][q] = p.in[k][base + q * 256];
#pragma unroll
    for (int k = 0; k < NW; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        __builtin_nontemporal_store(v[0][q] + (T)(k + 1) * v[1][q] - v[R - 1][q],
                                    p.out[k] + base + q * 256);
  }
}

template <class T>
static int run_step(long long images, int reps, int h) {
  const long long rows = images * 256;
  if (rows % h) {
    fprintf(stderr, "copy_calib: segment rows must divide %lld\n", rows);
    return 2;
  }
  const long long n = rows * 4 * 256;
  T* f[5] = {};
  for (auto& q : f) {
    if (hipMalloc(&q, sizeof(T) * n) != hipSuccess) return 1;
    if (hipMemset(q, 0, sizeof(T) * n) != hipSuccess) return 1;
  }
  T *x[2] = {f[0], f[1]}, *g[2] = {f[2], f[3]}, *d = f[4];
  const unsigned grid = (unsigned)(rows / h);
  hipEvent_t ev[3];
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return 1;
  double tg = 0, tp = 0;
  const int warm = 20;
  for (int it = 0; it < reps + warm; ++it) {
    const int i = it & 1;
    StepBufs<T, 3, 2> pg{{x[i], d, g[i]}, {x[1 - i], g[1 - i]}};
    StepBufs<T, 3, 1> pp{{g[1 - i], d, x[1 - i]}, {d}};
    if (hipEventRecord(ev[0], 0) != hipSuccess) return 1;
    hipLaunchKernelGGL((step_kernel<T, 3, 2>), dim3(grid), dim3(256), 0, 0, pg, h);
    if (hipEventRecord(ev[1], 0) != hipSuccess) return 1;
    hipLaunchKernelGGL((step_kernel<T, 3, 1>), dim3(grid), dim3(256), 0, 0, pp, h);
    if (hipEventRecord(ev[2], 0) != hipSuccess) return 1;
    if (hipEventSynchronize(ev[2]) != hipSuccess) return 1;
    float a = 0, c = 0;
    if (hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess) return 1;
    if (hipEventElapsedTime(&c, ev[1], ev[2]) != hipSuccess) return 1;
    if (it >= warm) {
      tg += a;
      tp += c;
    }
  }
  tg /= reps;
  tp /= reps;
  const double bytes = (double)sizeof(T) * n;
  printf("step %lld images, %zu bytes per element, %d-row segments: G-like %.2f us (%.1f GB/s), "
         "P-like %.2f us (%.1f GB/s)\n",
         images, sizeof(T), h, tg * 1e3, 5.0 * bytes / (tg * 1e-3) / 1e9, tp * 1e3,
         4.0 * bytes / (tp * 1e-3) / 1e9);
  for (auto* q : f) (void)hipFree(q);
  return 0;
}

template <int W>
static int run(long long mib, int reps, const char* mode, int waves) {
  using T = typename Elem<W>::T;
  const long long bytes = mib << 20;
  const long long n = bytes / (long long)sizeof(T);
  const bool stream = !strcmp(mode, "stream");
  const bool triad = !strcmp(mode, "triad") || stream;
  T *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&d, bytes) != hipSuccess) return 1;
  if (triad && (hipMalloc(&b, bytes) != hipSuccess || hipMalloc(&c, bytes) != hipSuccess)) return 1;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(d, 0, bytes) != hipSuccess) return 1;
  if (triad && (hipMemset(b, 0, bytes) != hipSuccess || hipMemset(c, 0, bytes) != hipSuccess))
    return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 1;
  const unsigned sblocks = (unsigned)(cus * waves);  // 4-wavefront blocks: `waves` per SIMD
  double moved = 0;
  auto launch = [&]() {
    if (!strcmp(mode, "read")) {
      hipLaunchKernelGGL(read_kernel<W>, dim3(blocks), dim3(256), 0, 0, a, (double*)d, n);
      moved = (double)bytes;
    } else if (!strcmp(mode, "write")) {
      hipLaunchKernelGGL(write_kernel<W>, dim3(blocks), dim3(256), 0, 0, d, n);
      moved = (double)bytes;
    } else if (stream) {
      hipLaunchKernelGGL(stream_kernel<W>, dim3(sblocks), dim3(256), 0, 0, a, b, c, d, n);
      moved = 4.0 * bytes;
    } else if (triad) {
      hipLaunchKernelGGL(triad_kernel<W>, dim3(blocks), dim3(256), 0, 0, a, b, c, d, n);
      moved = 4.0 * bytes;
    } else {
      hipLaunchKernelGGL(copy_kernel<W>, dim3(blocks), dim3(256), 0, 0, a, d, n);
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
  if (stream)
    printf("%s %lld MiB, %d bytes per lane, %d wavefronts per SIMD: %.4f ms per pass, %.1f GB/s\n",
           mode, mib, W, waves, per, moved / (per * 1e-3) / 1e9);
  else
    printf("%s %lld MiB, %d bytes per lane: %.4f ms per pass, %.1f GB/s\n", mode, mib, W, per,
           moved / (per * 1e-3) / 1e9);
  hipFree(a);
  hipFree(d);
  if (b) hipFree(b);
  if (c) hipFree(c);
  return 0;
}

int main(int argc, char** argv) {
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 20;
  const char* mode = argc > 3 ? argv[3] : "copy";
  const int width = argc > 4 ? atoi(argv[4]) : 16;
  const int waves = argc > 5 ? atoi(argv[5]) : 4;
  if (!strcmp(mode, "step")) {
    // mib: images; waves: segment rows
    const int h = argc > 5 ? atoi(argv[5]) : 16;
    if (mib <= 0 || mib > 256 || reps <= 0 || h < 1 || h > 256) {
      fprintf(stderr, "copy_calib step: 1..256 images, repeats > 0, 1..256 segment rows\n");
      return 2;
    }
    if (width == 4) return run_step<float>(mib, reps, h);
    if (width == 8) return run_step<double>(mib, reps, h);
    fprintf(stderr, "copy_calib step: bytes per element must be 4 or 8\n");
    return 2;
  }
  if (mib <= 0 || mib > 16384 || reps <= 0 || waves < 1 || waves > 8) {
    fprintf(stderr, "copy_calib: MiB in 1..16384, repeats > 0, 1..8 wavefronts per SIMD\n");
    return 2;
  }
  switch (width) {
    case 4: return run<4>(mib, reps, mode, waves);
    case 8: return run<8>(mib, reps, mode, waves);
    case 16: return run<16>(mib, reps, mode, waves);
    default:
      fprintf(stderr, "copy_calib: bytes per lane must be 4, 8 or 16\n");
      return 2;
  }
}
