// Local gauge alignment of neighbouring images (no reference function: the
// reference measures image distances with the gauge-invariant metric s_jb,
// mepgl.cu:31-32, and interpolates the raw fields).
//
// For image n >= 1, xi_n minimises |A_n + D xi - A_{n-1}|^2 over the grid,
// D being the forward difference (zero in the last column for x and the last
// row for y).  The normal equations L xi = dx^2 D^T (A_{n-1} - A_n), with L
// the Neumann graph Laplacian of the grid, are solved exactly in its
// eigenbasis, the orthonormal DCT-II C[i][k] = w_k cos(pi k (i + 1/2) / N)
// with eigenvalues lam_k = 2 - 2 cos(pi k / N):
//   xi = C ((C^T b C) / (lam_l + lam_k)) C^T,  the constant mode set to zero,
// as four dense products per image (k_dct_gemm).  The transform accumulates
// along the string, chi_n = chi_{n-1} + xi_n, and is applied as
// A_n += D chi_n and psi_a,n *= exp(i q_a chi_n) (fields, direction and both
// gradient buffers; the A parts of the direction and gradients do not
// change).  The forward curl of every plaquette is unchanged, and no
// neighbour A distance grows.  Strings with a fixed A (no self field) are
// left as they are.
//
// Work buffers: planes 0 and 1 of the string's double scratch in the fields'
// layout (cubic_rhs): b and the stage products, then xi and chi in plane 0.
#include "gauge_local.hpp"
#include "kernels_common.hpp"

namespace mepgl {

// Right-hand side b = dx^2 D^T (A_{n-1} - A_n) of image n (zero for image 0):
// a site's x-bond to its left and y-bond below enter with +R, its own bonds
// with -R (R = A_{n-1} - A_n, times dx for the dx^2 / dx scaling).
template <class T>
__global__ __launch_bounds__(kBlock) void k_glocal_rhs(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nsf) return;
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const int N = s.N, i = t.i;
  const long long IS = s.image_stride, PS = s.plane_stride, P = s.pitch;
  double* W = (double*)s.cubic_rhs + (long long)t.n * IS;
  const T* X1 = (const T*)s.x[a.xin] + (long long)t.n * IS;
  const T* X0 = X1 - IS;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long o = (long long)j * P + i;
    double v = 0;
    if (t.n > 0) {
      auto R = [&](int plane, long long off) {
        return (double)X0[plane * PS + off] - (double)X1[plane * PS + off];
      };
      if (i >= 1) v += R(PAX, o - 1);
      if (i <= N - 2) v -= R(PAX, o);
      if (j >= 1) v += R(PAY, o - P);
      if (j <= N - 2) v -= R(PAY, o);
    }
    W[o] = s.dx * v;
  }
}

// out(m, n) = sum_k A(m, k) B(k, n) for every image of one string (blockIdx.z),
// A(m, k) = A[m ar + k ac], B(k, n) = B[k br + n bc]; 64 x 64 output tiles,
// 16-deep k tiles in LDS, 4 x 4 outputs per thread.  With lam, the output is
// divided by lam[m] + lam[n] and its (0, 0) element set to zero.
constexpr int kGemmTile = 64, kGemmK = 16, kGemmThreads = 256;

__global__ __launch_bounds__(kGemmThreads) void k_dct_gemm(GemmArgs g) {
  __shared__ double As[kGemmK][kGemmTile + 1];
  __shared__ double Bs[kGemmK][kGemmTile + 1];
  const int N = g.n;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * kGemmTile, n0 = blockIdx.x * kGemmTile;
  const long long z = blockIdx.z;
  const double* A = g.A + z * g.abatch;
  const double* B = g.B + z * g.bbatch;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0;
  for (int k0 = 0; k0 < N; k0 += kGemmK) {
    // Consecutive threads read consecutive addresses of whichever operand
    // dimension is contiguous.
    for (int e = threadIdx.x; e < kGemmTile * kGemmK; e += kGemmThreads) {
      int mm, kk;
      if (g.ac == 1) {
        mm = e / kGemmK;
        kk = e % kGemmK;
      } else {
        mm = e % kGemmTile;
        kk = e / kGemmTile;
      }
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < N && k < N) ? A[m * g.ar + k * g.ac] : 0.0;
      int nn, kb;
      if (g.bc == 1) {
        nn = e % kGemmTile;
        kb = e / kGemmTile;
      } else {
        nn = e / kGemmK;
        kb = e % kGemmK;
      }
      const int n = n0 + nn, kB = k0 + kb;
      Bs[kb][nn] = (n < N && kB < N) ? B[kB * g.br + n * g.bc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGemmK; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(av[r], bv[c], acc[r][c]);
    }
    __syncthreads();
  }
  double* O = g.O + z * g.obatch;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = m0 + ty + 16 * r;
    if (m >= N) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int n = n0 + tx + 16 * c;
      if (n >= N) continue;
      double v = acc[r][c];
      if (g.lam) v = (m == 0 && n == 0) ? 0.0 : v / (g.lam[m] + g.lam[n]);
      O[m * g.orow + n] = v;
    }
  }
}

// chi_n = chi_{n-1} + xi_n in image order, one thread per site of one string.
__global__ __launch_bounds__(256) void k_glocal_prefix(double* W, int N, int F, long long pitch,
                                                      long long IS) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)N * N) return;
  const long long o = (e / N) * pitch + e % N;
  double acc = W[o];  // image 0: zero
  for (int n = 1; n < F; ++n) {
    acc = acc + W[(long long)n * IS + o];
    W[(long long)n * IS + o] = acc;
  }
}

// A_n += D chi_n; psi_a,n *= exp(i q_a chi_n) in the fields, the direction
// and both gradient buffers.
template <class T>
__global__ __launch_bounds__(kBlock) void k_glocal_apply(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nsf) return;
  const Strip t = strip_of(s, it);
  if (t.n == 0 || t.i >= s.N) return;
  const int N = s.N, i = t.i;
  const long long IS = s.image_stride, PS = s.plane_stride, P = s.pitch;
  const double* X = (const double*)s.cubic_rhs + (long long)t.n * IS;
  T* XA = (T*)s.x[a.xin] + (long long)t.n * IS;
  T* bufs[4] = {XA, (T*)s.d[a.dcur] + (long long)t.n * IS, (T*)s.g[0] + (long long)t.n * IS,
                (T*)s.g[1] + (long long)t.n * IS};
  for (int j = t.j0; j < t.j1; ++j) {
    const long long o = (long long)j * P + i;
    const double chi = X[o];
    if (i <= N - 2) XA[PAX * PS + o] = (T)((double)XA[PAX * PS + o] + (X[o + 1] - chi) / s.dx);
    if (j <= N - 2) XA[PAY * PS + o] = (T)((double)XA[PAY * PS + o] + (X[o + P] - chi) / s.dx);
    for (int c = 0; c < s.nc; ++c) {
      const double th = s.q[c] * chi, cs = cos(th), sn = sin(th);
      for (int b = 0; b < 4; ++b) {
        T* B = bufs[b];
        const long long pu = (PU1 + 2 * c) * PS + o, pv = (PV1 + 2 * c) * PS + o;
        const double u = B[pu], v = B[pv];
        B[pu] = (T)(u * cs - v * sn);
        B[pv] = (T)(u * sn + v * cs);
      }
    }
  }
}

hipError_t launch_glocal(const AuxArgs& a, int which, bool fp32, hipStream_t st) {
  if (a.nitems <= 0) return hipSuccess;
  const dim3 g((unsigned)a.nitems), b(kBlock);
  if (which == 0) {
    if (fp32)
      hipLaunchKernelGGL(k_glocal_rhs<float>, g, b, 0, st, a);
    else
      hipLaunchKernelGGL(k_glocal_rhs<double>, g, b, 0, st, a);
  } else {
    if (fp32)
      hipLaunchKernelGGL(k_glocal_apply<float>, g, b, 0, st, a);
    else
      hipLaunchKernelGGL(k_glocal_apply<double>, g, b, 0, st, a);
  }
  return hipGetLastError();
}

hipError_t launch_dct_gemm(const GemmArgs& g, int F, hipStream_t st) {
  const unsigned t = (unsigned)((g.n + kGemmTile - 1) / kGemmTile);
  hipLaunchKernelGGL(k_dct_gemm, dim3(t, t, (unsigned)F), dim3(kGemmThreads), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_glocal_prefix(double* W, int N, int F, long long pitch, long long IS,
                                hipStream_t st) {
  const long long sites = (long long)N * N;
  hipLaunchKernelGGL(k_glocal_prefix, dim3((unsigned)((sites + 255) / 256)), dim3(256), 0, st, W,
                     N, F, pitch, IS);
  return hipGetLastError();
}

}  // namespace mepgl
