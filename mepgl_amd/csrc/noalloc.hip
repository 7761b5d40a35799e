// computeFreeEnergy on caller-owned device arrays, allocating nothing: the
// per-tile partial sums go into a file-scope device array instead of device
// memory the library allocates.  For callers that cannot call hipMalloc
// (code under HIP graph capture, serving loops, harnesses that forbid it).
//
// Work decomposition and arithmetic are those of the context's energy pass
// (k_energy in kernels.hip): one wavefront per 64-column strip of a segment
// of `seg` rows of one image, each lane summing its column's site energies
// (site_energy, in double) with two-sum compensation, one partial per work
// item, and a fixed-order total per image.  With the same segment height the
// per-image energies equal the context's bit for bit.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../include/mepgl.h"
#include "kernels_common.hpp"

namespace mepgl {

// Partials: one (sum, compensation) pair per work item (image x tile).
constexpr int kNoAllocItems = 65536;
static __device__ __align__(256) double scratch[2 * kNoAllocItems];

constexpr int kNoAllocMaxDevices = 64;
static double* g_scratch[kNoAllocMaxDevices];  // device address of `scratch` per device
static std::mutex g_scratch_mu;

// Neighbour bits of site (i, j) of a domain (buildNearestNeighbourMap,
// nlcg.cu:48-103): bit k set when neighbour k (0 +x, 1 +x+y, 2 +y, 3 -x+y,
// 4 -x, 5 -x-y, 6 -y, 7 +x-y) is in the domain; 0 outside it.
__device__ __forceinline__ int na_bits(const int32_t* dom, int i, int j, int N) {
  const long long idx = (long long)j * N + i;
  if (dom[idx] == 0) return 0;
  const bool xr = i != N - 1, xl = i != 0, yu = j != N - 1, yd = j != 0;
  int m = 0;
  if (xr && dom[idx + 1]) m |= 1 << 0;
  if (xr && yu && dom[idx + 1 + N]) m |= 1 << 1;
  if (yu && dom[idx + N]) m |= 1 << 2;
  if (yu && xl && dom[idx - 1 + N]) m |= 1 << 3;
  if (xl && dom[idx - 1]) m |= 1 << 4;
  if (xl && yd && dom[idx - 1 - N]) m |= 1 << 5;
  if (yd && dom[idx - N]) m |= 1 << 6;
  if (yd && xr && dom[idx - N + 1]) m |= 1 << 7;
  return m;
}

struct NaArgs {
  int N, F, tiles_x, tiles_y, seg, nsf;
  double inv_dx, dx2, q[2], eta, gamma, delta;
  DevParam a[2], b[2], mxx[2], myy[2], h;  // planes: device, [j][i], row length N
  const double* A;                         // (F, N, N, 2): A_x, A_y
  const double* P[2];                      // (F, N, N, 2): real, imaginary
  const int32_t* comp;                     // (N, N)
  const int32_t* sc;
  double* part;                            // the scratch array's device address
  double* out;                             // F energies (device)
};

template <int NC>
__global__ __launch_bounds__(kBlock) void k_energy_noalloc(NaArgs a) {
  const int ntiles = a.tiles_x * a.tiles_y;
  const int n = blockIdx.x / ntiles, tile = blockIdx.x % ntiles;
  const int N = a.N;
  const int i = (tile % a.tiles_x) * kTileX + (int)threadIdx.x;
  const int j0 = (tile / a.tiles_x) * a.seg, j1 = min(j0 + a.seg, N);
  const long long img = (long long)n * N * N;
  double acc_s[1] = {0}, acc_c[1] = {0};
  if (i < N) {
    const int ip = min(i + 1, N - 1), im = max(i - 1, 0);
    for (int j = j0; j < j1; ++j) {
      const int jp = min(j + 1, N - 1), jm = max(j - 1, 0);
      const long long own = (long long)j * N + i;
      // Slots 0 own, 1 +x, 2 -x, 3 +y, 4 -y (neighbours clamped to the grid).
      const long long o[5] = {own, (long long)j * N + ip, (long long)j * N + im,
                              (long long)jp * N + i, (long long)jm * N + i};
      PolyNbhd<double, NC> nb;
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        nb.ax[t] = a.A[2 * (img + o[t])];
        nb.ay[t] = a.A[2 * (img + o[t]) + 1];
        nb.dax[t] = nb.day[t] = 0;
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double* p = a.P[c];
        nb.u[c][0] = p[2 * (img + o[0])];
        nb.v[c][0] = p[2 * (img + o[0]) + 1];
        nb.u[c][1] = p[2 * (img + o[1])];
        nb.v[c][1] = p[2 * (img + o[1]) + 1];
        nb.u[c][2] = p[2 * (img + o[3])];
        nb.v[c][2] = p[2 * (img + o[3]) + 1];
#pragma unroll
        for (int t = 0; t < 3; ++t) nb.du[c][t] = nb.dv[c][t] = 0;
      }
      nb.h0 = a.h.p ? a.h.p[own] : a.h.v;
      CompConst<double> k[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double av = param_at<double>(a.a[c], own), bv = param_at<double>(a.b[c], own);
        const double mx = param_at<double>(a.mxx[c], own), my = param_at<double>(a.myy[c], own);
        k[c].q = a.q[c];
        k[c].two_a = 2.0 * av;
        k[c].two_b = 2.0 * bv;
        k[c].b = bv;
        k[c].ab = av / bv;
        k[c].hx = 0.5 / mx;
        k[c].hy = 0.5 / my;
      }
      const int scb = na_bits(a.sc, i, j, N), cmb = na_bits(a.comp, i, j, N);
      const double e =
          site_energy<NC>(nb, k, a.eta, a.gamma, a.delta, a.inv_dx, a.dx2, scb, cmb, a.nsf);
      two_sum_acc(acc_s[0], acc_c[0], e);
    }
  }
  publish_partial<1, true>(acc_s, acc_c, a.part, n, tile, ntiles);
}

// Fixed-order total of each image's partials (one wavefront per image).
__global__ __launch_bounds__(kBlock) void k_energy_noalloc_total(NaArgs a) {
  double fin[1];
  reduce_partials<1, true>(a.part, blockIdx.x, a.tiles_x * a.tiles_y, fin);
  if (threadIdx.x == 0) a.out[blockIdx.x] = fin[0];
}

}  // namespace mepgl

using namespace mepgl;

extern "C" int mepgl_free_energy_noalloc(const mepgl_string_desc* d, const double* a,
                                         const double* psi_1, const double* psi_2,
                                         double* energy, void* stream) {
  if (!d || !a || !psi_1 || !energy || !d->comp_domain || !d->sc_domain) return MEPGL_ERR_ARG;
  if (d->N < 2 || d->N > 65536 || d->F < 1 || d->F > kMaxImages ||
      (d->ncomp != 1 && d->ncomp != 2) || (d->ncomp == 2 && !psi_2) || !(d->dx > 0))
    return MEPGL_ERR_ARG;
  const mepgl_param* prm[] = {&d->a_1, &d->b_1, &d->m_xx_1, &d->m_yy_1, &d->a_2,
                              &d->b_2, &d->m_xx_2, &d->m_yy_2, &d->h};
  for (auto* p : prm)
    if (!p->plane && !std::isfinite(p->value)) return MEPGL_ERR_ARG;
  NaArgs x{};
  x.N = d->N;
  x.F = d->F;
  x.nsf = d->no_self_field ? 1 : 0;
  x.seg = d->segment_rows > 0 ? std::min(d->segment_rows, d->N) : std::min(16, d->N);
  x.tiles_x = (d->N + kTileX - 1) / kTileX;
  x.tiles_y = (d->N + x.seg - 1) / x.seg;
  // The partials of every work item must fit the scratch array.
  if ((long long)x.F * x.tiles_x * x.tiles_y > kNoAllocItems) return MEPGL_ERR_ARG;
  x.inv_dx = 1.0 / d->dx;
  x.dx2 = d->dx * d->dx;
  x.q[0] = d->q_1;
  x.q[1] = d->q_2;
  x.eta = d->eta;
  x.gamma = d->gamma;
  x.delta = d->delta;
  auto dp = [](const mepgl_param& p) { return DevParam{p.value, p.plane}; };
  x.a[0] = dp(d->a_1), x.b[0] = dp(d->b_1), x.mxx[0] = dp(d->m_xx_1), x.myy[0] = dp(d->m_yy_1);
  x.a[1] = dp(d->a_2), x.b[1] = dp(d->b_2), x.mxx[1] = dp(d->m_xx_2), x.myy[1] = dp(d->m_yy_2);
  x.h = dp(d->h);
  x.A = a;
  x.P[0] = psi_1;
  x.P[1] = d->ncomp == 2 ? psi_2 : psi_1;
  x.comp = d->comp_domain;
  x.sc = d->sc_domain;
  x.out = energy;
  // The device address of this device's scratch array, looked up once per
  // device with that device current (the host-side name is not a device
  // address and is never passed to a kernel).
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
    return MEPGL_ERR_HIP;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
      void* ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
        return MEPGL_ERR_HIP;
      g_scratch[dev] = (double*)ptr;
    }
    x.part = g_scratch[dev];
  }
  hipStream_t st = (hipStream_t)stream;
  const unsigned items = (unsigned)(x.F * x.tiles_x * x.tiles_y);
  if (d->ncomp == 2)
    hipLaunchKernelGGL(k_energy_noalloc<2>, dim3(items), dim3(kBlock), 0, st, x);
  else
    hipLaunchKernelGGL(k_energy_noalloc<1>, dim3(items), dim3(kBlock), 0, st, x);
  if (hipGetLastError() != hipSuccess) return MEPGL_ERR_HIP;
  hipLaunchKernelGGL(k_energy_noalloc_total, dim3((unsigned)x.F), dim3(kBlock), 0, st, x);
  return hipGetLastError() == hipSuccess ? MEPGL_OK : MEPGL_ERR_HIP;
}
