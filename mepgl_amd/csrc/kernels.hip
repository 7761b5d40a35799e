// HIP kernels of the MI355X gauged-string-method relaxation loop (gfx950).
//
// One NLCG step (nlcgSteps, nlcg.cu:1478-1583) is two launches over a work
// list of (string, image, tile) items:
//   G  fused pending update x += alpha d (updateFields), gradient
//      (gradientPsi + gradientA), per-workgroup partials of |g|^2 and
//      g.(g - g_old) (gradientDensity); the last workgroup of each image
//      reduces the partials in a fixed order and forms the Polak-Ribiere beta
//      (computeStepDirection's scalar part).
//   P  direction d = -g + beta d (computeStepDirection) evaluated on the
//      stencil, line-search quartic partials (polynomialExpansion); the last
//      workgroup of each image reduces them and runs the line search
//      (lineSearch) as a parallel scan over a host-built table.
// Reductions are deterministic (fixed tile order, fixed trees), with two-sum
// compensation in the mixed-precision path.  The arrival counters reset
// themselves, so launches can be replayed from a hipGraph.
#include <float.h>

#include "common.hpp"
#include "kernels.hpp"
#include "stencil.hpp"

namespace mepgl {

// ---------------------------------------------------------------------------
// Work-item mapping: consecutive items go to the same XCD so that the halo
// rows a tile shares with its neighbours are served by that XCD's L2.
__device__ __forceinline__ int xcd_item(int b, int nb) {
  const int q = nb / kNumXcd, r = nb % kNumXcd;
  const int k = b % kNumXcd, l = b / kNumXcd;
  return k * q + (k < r ? k : r) + l;
}

template <class T>
__device__ __forceinline__ T ldg(const T* p, long long off) {
  return p[off];
}

__device__ __forceinline__ double ld_coherent(const double* p) {
  unsigned long long v =
      __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)v);
}

// Two-sum accumulation (Knuth), exact error term.
__device__ __forceinline__ void two_sum_acc(double& s, double& c, double v) {
  const double t = s + v;
  const double bp = t - s;
  const double e = (s - (t - bp)) + (v - bp);
  s = t;
  c += e;
}

template <bool COMP>
__device__ __forceinline__ void acc(double& s, double& c, double v) {
  if (COMP)
    two_sum_acc(s, c, v);
  else
    s += v;
}

// Block-wide reduction of K (sum, compensation) pairs; thread 0 ends with
// the result.  lds must hold kWavesPerBlock * 2 * K doubles.
template <int K, bool COMP>
__device__ __forceinline__ void block_reduce(double (&s)[K], double (&c)[K], double* lds) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double so = __shfl_xor(s[k], off);
      if (COMP) {
        const double co = __shfl_xor(c[k], off);
        two_sum_acc(s[k], c[k], so);
        c[k] += co;
      } else {
        s[k] += so;
      }
    }
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      lds[w * 2 * K + k] = s[k];
      lds[w * 2 * K + K + k] = c[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double ss = lds[k], cc = lds[K + k];
      for (int ww = 1; ww < kWavesPerBlock; ++ww) {
        if (COMP) {
          two_sum_acc(ss, cc, lds[ww * 2 * K + k]);
          cc += lds[ww * 2 * K + K + k];
        } else {
          ss += lds[ww * 2 * K + k];
        }
      }
      s[k] = ss;
      c[k] = cc;
    }
  }
  __syncthreads();
}

// Writes this workgroup's K pairs to part[(n * ntiles + tile) * 2K], then
// returns true in every thread of the last workgroup of image n to arrive,
// with the fixed-order reduction over all tiles in fin[K] (valid in every
// thread of that workgroup).  The counter is reset by the last workgroup.
template <int K, bool COMP>
__device__ bool finish_image(double (&s)[K], double (&c)[K], double* lds, double* part,
                             unsigned* count, int n, int tile, int ntiles, double (&fin)[K]) {
  __shared__ int s_last;
  block_reduce<K, COMP>(s, c, lds);
  double* mine = part + ((long long)n * ntiles + tile) * (2 * K);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      __hip_atomic_store(&mine[k], s[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&mine[K + k], c[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned prev =
        __hip_atomic_fetch_add(&count[n], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == (unsigned)(ntiles - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double ss[K], cc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) ss[k] = cc[k] = 0.0;
  const double* base = part + (long long)n * ntiles * (2 * K);
  for (int t = threadIdx.x; t < ntiles; t += kBlock) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      acc<COMP>(ss[k], cc[k], ld_coherent(&base[(long long)t * 2 * K + k]));
      if (COMP) cc[k] += ld_coherent(&base[(long long)t * 2 * K + K + k]);
    }
  }
  block_reduce<K, COMP>(ss, cc, lds);
  __shared__ double s_fin[K];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s_fin[k] = ss[k] + cc[k];
    __hip_atomic_store(&count[n], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) fin[k] = s_fin[k];
  return true;
}

// ---------------------------------------------------------------------------
// Tile geometry of a work item.
struct TileCtx {
  int n, tile, i, j0, rstride, nrows;
};

__device__ __forceinline__ TileCtx tile_of(const DevString& s, const WorkItem& it) {
  TileCtx t;
  t.n = it.image;
  t.tile = it.tile;
  const int tx = it.tile % s.tiles_x, ty = it.tile / s.tiles_x;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  t.i = tx * kTileX + lane;
  t.j0 = ty * (kWavesPerBlock * s.rows_per_thread) + w;
  t.rstride = kWavesPerBlock;
  t.nrows = s.rows_per_thread;
  return t;
}

__device__ __forceinline__ double hval(const DevString& s, long long off) {
  return s.h.p ? s.h.p[off] : s.h.v;
}

// ---------------------------------------------------------------------------
// G kernel body.
template <class T, int NC>
__device__ void grad_body(const StepArgs& a, const DevString& s, const WorkItem& it,
                          double* lds) {
  constexpr bool COMP = sizeof(T) == 4;
  const TileCtx tc = tile_of(s, it);
  const int N = s.N, P = s.pitch;
  const long long ps = s.plane_stride;
  const long long ib = (long long)tc.n * s.image_stride;
  const T* X = (const T*)s.x[a.xin] + ib;
  T* XO = (T*)s.x[a.xout] + ib;
  const T* D = (const T*)s.d[a.din] + ib;
  T* G = (T*)(a.scratch ? s.gs : s.g[a.gcur]) + ib;
  const T* GO = (const T*)s.g[a.gold] + ib;
  ImageScalars& sc = a.sc[s.img0 + tc.n];
  const bool upd = a.upd;
  const T al = upd ? (T)sc.alpha : (T)0;
  const bool do_a = !s.nsf;
  const bool do_psi = !a.maxwell;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};

  auto fld = [&](int p, long long off) -> T {
    const T xv = X[p * ps + off];
    return upd ? fma(al, D[p * ps + off], xv) : xv;
  };

  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= N || j >= N) continue;
    const long long off = (long long)j * P + i;
    const int ip = min(i + 1, N - 1), im = max(i - 1, 0);
    const int jp = min(j + 1, N - 1), jm = max(j - 1, 0);
    const long long offs[5] = {off, (long long)j * P + ip, (long long)j * P + im,
                               (long long)jp * P + i, (long long)jm * P + i};
    const uint16_t nnv = s.nn[off];
    const int sc_bits = nnv & 0xff, cm_bits = nnv >> 8;

    Nbhd<T, NC> nb;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int jj = dy == 0 ? jm : (dy == 1 ? j : jp);
#pragma unroll
      for (int dxx = 0; dxx < 3; ++dxx) {
        const int ii = dxx == 0 ? im : (dxx == 1 ? i : ip);
        const long long o = (long long)jj * P + ii;
        const bool need = do_a || (dy == 1) || (dxx == 1);
        nb.ax[dy][dxx] = need ? fld(PAX, o) : (T)0;
        nb.ay[dy][dxx] = need ? fld(PAY, o) : (T)0;
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        nb.u[c][t] = fld(PU1 + 2 * c, offs[t]);
        nb.v[c][t] = fld(PV1 + 2 * c, offs[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 5; ++t) nb.h[t] = (T)hval(s, offs[t]);

    CompConst<T> k[NC];
    load_comp_const<T, NC>(s, off, k);
    T gp[NC][2], ga[2] = {0, 0};
    site_gradient<T, NC>(nb, k, eta, gam, del, idx, dx2, sc_bits, cm_bits, do_psi, do_a, gp, ga);

    // Stores: gradient, then the updated fields of this site.
    if (do_psi) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        G[(PU1 + 2 * c) * ps + off] = gp[c][0];
        G[(PV1 + 2 * c) * ps + off] = gp[c][1];
      }
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        gp[c][0] = G[(PU1 + 2 * c) * ps + off];
        gp[c][1] = G[(PV1 + 2 * c) * ps + off];
      }
    }
    if (do_a) {
      G[PAX * ps + off] = ga[0];
      G[PAY * ps + off] = ga[1];
    }
    if (upd) {
      if (do_a) {
        XO[PAX * ps + off] = nb.ax[1][1];
        XO[PAY * ps + off] = nb.ay[1][1];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        XO[(PU1 + 2 * c) * ps + off] = nb.u[c][0];
        XO[(PV1 + 2 * c) * ps + off] = nb.v[c][0];
      }
    }
    if (!a.scratch) {
      // gradientDensity, nlcg.cu:919-949
      double n2 = 0, dg = 0;
      if (do_a) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const double gv = (double)ga[q], go = (double)GO[q * ps + off];
          n2 += gv * gv;
          dg += gv * (gv - go);
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const double gv = (double)gp[c][q], go = (double)GO[(PU1 + 2 * c + q) * ps + off];
          n2 += gv * gv;
          dg += gv * (gv - go);
        }
      }
      acc<COMP>(acc_s[0], acc_c[0], n2);
      acc<COMP>(acc_s[1], acc_c[1], dg);
    }
  }
  if (a.scratch) return;

  double fin[2];
  const int ntiles = s.tiles_x * s.tiles_y;
  if (finish_image<2, COMP>(acc_s, acc_c, lds, s.part_g, s.count_g, tc.n, tc.tile, ntiles, fin)) {
    if (threadIdx.x == 0) {
      sc.g_norm2[a.scur] = fin[0];
      sc.g_delta_g = fin[1];
      double beta = 0;
      if (a.nlcg) {
        beta = fin[1] / sc.g_norm2[a.sold];
        beta = fmax(beta, 0.0);
      }
      sc.beta = beta;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_grad(StepArgs a) {
  __shared__ double lds[kWavesPerBlock * 2 * kPartP];
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    grad_body<T, 2>(a, s, it, lds);
  else
    grad_body<T, 1>(a, s, it, lds);
}

// ---------------------------------------------------------------------------
// Line search, lineSearch nlcg.cu:1398-1445, as a parallel scan: candidate p
// compares f(x_{p+1}) with f(x_p); the first rise ends the scan and
// alpha = x_{p+1} / m.  f is evaluated without contraction, in the
// reference's operation order, so the result equals the sequential loop's.
__device__ __forceinline__ double quartic_at(const LineSearchTable& t, int k, double A, double B,
                                             double C, double D, double E) {
  const double x = t.x[k];
  double f = __dmul_rn(E, t.x4[k]);
  f = __dadd_rn(f, __dmul_rn(__dmul_rn(__dmul_rn(D, x), x), x));
  f = __dadd_rn(f, __dmul_rn(__dmul_rn(C, x), x));
  f = __dadd_rn(f, __dmul_rn(B, x));
  f = __dadd_rn(f, A);
  return f;
}

__device__ int line_search_block(const LineSearchTable& t, double A, double B, double C, double D,
                                 double E) {
  __shared__ int s_first;
  for (int base = 0; base < kLsSteps; base += kBlock) {
    if (threadIdx.x == 0) s_first = kLsSteps;
    __syncthreads();
    const int p = base + threadIdx.x;
    if (p < kLsSteps) {
      const double f_old = quartic_at(t, p, A, B, C, D, E);
      const double f_new = quartic_at(t, p + 1, A, B, C, D, E);
      if (f_new > f_old) atomicMin(&s_first, p);
    }
    __syncthreads();
    const int first = s_first;
    __syncthreads();
    if (first < kLsSteps) return first;
  }
  return kLsSteps - 1;
}

// P kernel body.
template <class T, int NC>
__device__ void poly_body(const StepArgs& a, const DevString& s, const WorkItem& it,
                          double* lds) {
  constexpr bool COMP = sizeof(T) == 4;
  const TileCtx tc = tile_of(s, it);
  const int N = s.N, P = s.pitch;
  const long long ps = s.plane_stride;
  const long long ib = (long long)tc.n * s.image_stride;
  const T* X = (const T*)s.x[a.xin] + ib;
  const T* G = (const T*)s.g[a.gcur] + ib;
  const T* D = (const T*)s.d[a.din] + ib;
  T* DO = (T*)s.d[a.dout] + ib;
  ImageScalars& sc = a.sc[s.img0 + tc.n];
  const T beta = (T)sc.beta;
  const bool nsf = s.nsf;
  const bool exact = s.poly_exact;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  double acc_s[5] = {0, 0, 0, 0, 0}, acc_c[5] = {0, 0, 0, 0, 0};

  // d = -g + beta d  (computeStepDirection, nlcg.cu:985-993)
  auto dir = [&](int p, long long off) -> T { return fma(beta, D[p * ps + off], -G[p * ps + off]); };

  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= N || j >= N) continue;
    const long long off = (long long)j * P + i;
    const int ip = min(i + 1, N - 1), im = max(i - 1, 0);
    const int jp = min(j + 1, N - 1), jm = max(j - 1, 0);
    const long long offs[5] = {off, (long long)j * P + ip, (long long)j * P + im,
                               (long long)jp * P + i, (long long)jm * P + i};
    const uint16_t nnv = s.nn[off];
    const int sc_bits = nnv & 0xff, cm_bits = nnv >> 8;

    PolyNbhd<T, NC> nb;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      nb.ax[t] = X[PAX * ps + offs[t]];
      nb.ay[t] = X[PAY * ps + offs[t]];
      nb.dax[t] = nsf ? (T)0 : dir(PAX, offs[t]);
      nb.day[t] = nsf ? (T)0 : dir(PAY, offs[t]);
    }
    const int ps3[3] = {0, 1, 3};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const long long o = offs[ps3[t]];
        nb.u[c][t] = X[(PU1 + 2 * c) * ps + o];
        nb.v[c][t] = X[(PV1 + 2 * c) * ps + o];
        nb.du[c][t] = dir(PU1 + 2 * c, o);
        nb.dv[c][t] = dir(PV1 + 2 * c, o);
      }
    }
    nb.h0 = (T)hval(s, off);

    if (!nsf) {
      DO[PAX * ps + off] = nb.dax[0];
      DO[PAY * ps + off] = nb.day[0];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      DO[(PU1 + 2 * c) * ps + off] = nb.du[c][0];
      DO[(PV1 + 2 * c) * ps + off] = nb.dv[c][0];
    }

    CompConst<T> k[NC];
    load_comp_const<T, NC>(s, off, k);
    const Quartic<T> q = site_poly<T, NC>(nb, k, eta, gam, del, idx, dx2, sc_bits, cm_bits, nsf,
                                          exact);
#pragma unroll
    for (int t = 0; t < 5; ++t) acc<COMP>(acc_s[t], acc_c[t], (double)q.c[t]);
  }

  double fin[5];
  const int ntiles = s.tiles_x * s.tiles_y;
  if (finish_image<5, COMP>(acc_s, acc_c, lds, s.part_p, s.count_p, tc.n, tc.tile, ntiles, fin)) {
    const int first = line_search_block(*a.ls, fin[0], fin[1], fin[2], fin[3], fin[4]);
    if (threadIdx.x == 0) {
      sc.fenergy = fin[0];
      sc.poly[0] = fin[1];
      sc.poly[1] = fin[2];
      sc.poly[2] = fin[3];
      sc.poly[3] = fin[4];
      sc.alpha = a.ls->alpha[first];
      sc.ls_index = first;
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_poly(StepArgs a) {
  __shared__ double lds[kWavesPerBlock * 2 * kPartP];
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    poly_body<T, 2>(a, s, it, lds);
  else
    poly_body<T, 1>(a, s, it, lds);
}

// ---------------------------------------------------------------------------
// Pending update applied in place (updateFields, nlcg.cu:1372-1396), used
// before anything other than the next gradient reads the fields.
template <class T>
__global__ __launch_bounds__(kBlock) void k_update(StepArgs a) {
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  const TileCtx tc = tile_of(s, it);
  const long long ib = (long long)tc.n * s.image_stride;
  T* X = (T*)s.x[a.xin] + ib;
  const T* D = (const T*)s.d[a.din] + ib;
  const T al = (T)a.sc[s.img0 + tc.n].alpha;
  const int p0 = s.nsf ? 2 : 0;
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    for (int p = p0; p < s.nplanes; ++p) {
      const long long o = p * s.plane_stride + off;
      X[o] = fma(al, D[o], X[o]);
    }
  }
}

// ---------------------------------------------------------------------------
// Shared gather of the 5-point neighbourhood in double (energy, JB).
template <class T, int NC>
__device__ __forceinline__ void gather5(const DevString& s, const T* X, int i, int j,
                                        PolyNbhd<double, NC>& nb, double (&um)[NC],
                                        double (&vm)[NC], double (&uym)[NC], double (&vym)[NC]) {
  const int N = s.N, P = s.pitch;
  const long long ps = s.plane_stride;
  const int ip = min(i + 1, N - 1), im = max(i - 1, 0);
  const int jp = min(j + 1, N - 1), jm = max(j - 1, 0);
  const long long off = (long long)j * P + i;
  const long long offs[5] = {off, (long long)j * P + ip, (long long)j * P + im,
                             (long long)jp * P + i, (long long)jm * P + i};
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    nb.ax[t] = (double)X[PAX * ps + offs[t]];
    nb.ay[t] = (double)X[PAY * ps + offs[t]];
    nb.dax[t] = nb.day[t] = 0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    nb.u[c][0] = (double)X[(PU1 + 2 * c) * ps + offs[0]];
    nb.v[c][0] = (double)X[(PV1 + 2 * c) * ps + offs[0]];
    nb.u[c][1] = (double)X[(PU1 + 2 * c) * ps + offs[1]];
    nb.v[c][1] = (double)X[(PV1 + 2 * c) * ps + offs[1]];
    nb.u[c][2] = (double)X[(PU1 + 2 * c) * ps + offs[3]];
    nb.v[c][2] = (double)X[(PV1 + 2 * c) * ps + offs[3]];
    um[c] = (double)X[(PU1 + 2 * c) * ps + offs[2]];
    vm[c] = (double)X[(PV1 + 2 * c) * ps + offs[2]];
    uym[c] = (double)X[(PU1 + 2 * c) * ps + offs[4]];
    vym[c] = (double)X[(PV1 + 2 * c) * ps + offs[4]];
#pragma unroll
    for (int t = 0; t < 3; ++t) nb.du[c][t] = nb.dv[c][t] = 0;
  }
  nb.h0 = hval(s, off);
}

// computeFreeEnergy (nlcg.cu:105-240), always evaluated in double.
template <class T, int NC>
__device__ void energy_body(const AuxArgs& a, const DevString& s, const WorkItem& it,
                            double* lds) {
  const TileCtx tc = tile_of(s, it);
  const T* X = (const T*)s.x[a.xin] + (long long)tc.n * s.image_stride;
  const double idx = s.inv_dx, dx2 = s.dx2;
  double acc_s[1] = {0}, acc_c[1] = {0};
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    PolyNbhd<double, NC> nb;
    double um[NC], vm[NC], uym[NC], vym[NC];
    gather5<T, NC>(s, X, i, j, nb, um, vm, uym, vym);
    CompConst<double> k[NC];
    load_comp_const<double, NC>(s, off, k);
    const uint16_t nnv = s.nn[off];
    const double e = site_energy<NC>(nb, k, s.eta, s.gamma, s.delta, idx, dx2, nnv & 0xff,
                                     nnv >> 8, s.nsf);
    two_sum_acc(acc_s[0], acc_c[0], e);
  }
  double fin[1];
  if (finish_image<1, true>(acc_s, acc_c, lds, s.part_p, s.count_p, tc.n, tc.tile,
                            s.tiles_x * s.tiles_y, fin)) {
    if (threadIdx.x == 0) a.out[s.img0 + tc.n] = fin[0];
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_energy(AuxArgs a) {
  __shared__ double lds[kWavesPerBlock * 2 * kPartP];
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    energy_body<T, 2>(a, s, it, lds);
  else
    energy_body<T, 1>(a, s, it, lds);
}

// computeJB (nlcg.cu:243-380) into planes b, j_1, j_2 of a double buffer
// laid out [image][plane][row][pitch] with 1 + 2*nc planes.
template <class T, int NC>
__device__ void jb_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const TileCtx tc = tile_of(s, it);
  const T* X = (const T*)s.x[a.xin] + (long long)tc.n * s.image_stride;
  double* O = a.jb[it.string] + (long long)tc.n * (1 + 2 * NC) * s.plane_stride;
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    PolyNbhd<double, NC> nb;
    double um[NC], vm[NC], uym[NC], vym[NC];
    gather5<T, NC>(s, X, i, j, nb, um, vm, uym, vym);
    CompConst<double> k[NC];
    load_comp_const<double, NC>(s, off, k);
    const uint16_t nnv = s.nn[off];
    double b, jj[NC][2];
    site_jb<NC>(nb, um, vm, uym, vym, k, s.inv_dx, nnv & 0xff, nnv >> 8, b, jj);
    O[off] = b;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      O[(1 + 2 * c) * s.plane_stride + off] = jj[c][0];
      O[(2 + 2 * c) * s.plane_stride + off] = jj[c][1];
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_jb(AuxArgs a) {
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    jb_body<T, 2>(a, s, it);
  else
    jb_body<T, 1>(a, s, it);
}

// Arclength partials (arcLengthParametrization + arcLengthParametrizationJB,
// interpolation.cu:39-92): squared distance of image n to image n-1 in the
// plain metric and in the gauge-invariant (b, j) metric, with b and j
// recomputed on the fly for both images.  Images 0 contribute nothing.
template <class T, int NC>
__device__ void arclen_body(const AuxArgs& a, const DevString& s, const WorkItem& it,
                            double* lds) {
  const TileCtx tc = tile_of(s, it);
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  if (tc.n > 0) {
    const T* X1 = (const T*)s.x[a.xin] + (long long)tc.n * s.image_stride;
    const T* X0 = X1 - s.image_stride;
    for (int r = 0; r < tc.nrows; ++r) {
      const int i = tc.i, j = tc.j0 + r * tc.rstride;
      if (i >= s.N || j >= s.N) continue;
      const long long off = (long long)j * s.pitch + i;
      double d2 = 0;
      for (int p = 0; p < 2 + 2 * NC; ++p) {
        const double dv = (double)X1[p * s.plane_stride + off] - (double)X0[p * s.plane_stride + off];
        d2 += dv * dv;
      }
      CompConst<double> k[NC];
      load_comp_const<double, NC>(s, off, k);
      const uint16_t nnv = s.nn[off];
      double b[2], jj[2][NC][2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        PolyNbhd<double, NC> nb;
        double um[NC], vm[NC], uym[NC], vym[NC];
        gather5<T, NC>(s, m ? X1 : X0, i, j, nb, um, vm, uym, vym);
        site_jb<NC>(nb, um, vm, uym, vym, k, s.inv_dx, nnv & 0xff, nnv >> 8, b[m], jj[m]);
      }
      double djb = sq(b[1] - b[0]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
        djb += sq(jj[1][c][0] - jj[0][c][0]) + sq(jj[1][c][1] - jj[0][c][1]);
      two_sum_acc(acc_s[0], acc_c[0], d2);
      two_sum_acc(acc_s[1], acc_c[1], djb);
    }
  }
  double fin[2];
  if (finish_image<2, true>(acc_s, acc_c, lds, s.part_a, s.count_g, tc.n, tc.tile,
                            s.tiles_x * s.tiles_y, fin)) {
    if (threadIdx.x == 0) {
      ImageScalars& sc = a.sc[s.img0 + tc.n];
      sc.ds2 = fin[0];
      sc.ds2_jb = fin[1];
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_arclen(AuxArgs a) {
  __shared__ double lds[kWavesPerBlock * 2 * kPartP];
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    arclen_body<T, 2>(a, s, it, lds);
  else
    arclen_body<T, 1>(a, s, it, lds);
}

// computeArcLength (interpolation.cu:126-140) for both metrics, then the
// redistribution plan of every new image on s_jb: the segment search of
// generateNewFramesLinear (interpolation.cu:200-219) and of
// generateNewFramesCubic (interpolation.cu:373-390), and the Thomas factors
// of the spline system built by buildHMatrix (interpolation.cu:252-290).
// One thread per string.
__global__ void k_arc_finalize(AuxArgs a, int nstrings) {
  const int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  sc[0].s = 0;
  sc[0].s_jb = 0;
  for (int n = 1; n < F; ++n) {
    sc[n].s = sc[n - 1].s + sqrt(sc[n].ds2);
    sc[n].s_jb = sc[n - 1].s_jb + sqrt(sc[n].ds2_jb);
  }
  for (int n = 0; n < F; ++n) {
    sc[n].s /= sc[F - 1].s;
    sc[n].s_jb /= sc[F - 1].s_jb;
  }
  // Linear plan (F_new = F).
  {
    int k = 1;
    double seg_start = 0, seg_end = sc[1].s_jb;
    for (int n = 0; n < F; ++n) {
      const double s_new = (double)n / ((double)F - 1);
      while (seg_end < s_new && k + 1 < F) {
        ++k;
        seg_start = sc[k - 1].s_jb;
        seg_end = sc[k].s_jb;
      }
      s.plan_k[n] = k;
      s.plan_w[n] = (s_new - seg_start) / (sc[k].s_jb - sc[k - 1].s_jb);
    }
  }
  // Cubic plan and spline factors (need F >= 4).
  if (F >= 4) {
    double* cb = s.cubic;  // [F][4]: c', m, dl, h
    for (int n = 0; n < F; ++n) cb[4 * n + 3] = n == 0 ? 0.0 : sc[n].s_jb - sc[n - 1].s_jb;
    auto h = [&](int n) { return cb[4 * n + 3]; };
    double dl[2], d, du;
    for (int n = 0; n < F; ++n) {
      if (n == 0) {
        dl[0] = 0; d = -h(1); du = h(1);
      } else if (n == F - 1) {
        dl[0] = h(F - 1); d = -h(F - 1); du = 0;
      } else {
        dl[0] = h(n); d = 2 * (h(n) + h(n + 1)); du = h(n + 1);
      }
      const double m = n == 0 ? d : d - dl[0] * cb[4 * (n - 1) + 0];
      cb[4 * n + 0] = du / m;
      cb[4 * n + 1] = m;
      cb[4 * n + 2] = dl[0];
    }
    int k = 1;
    double seg_start = 0, seg_end = h(1);
    for (int n = 0; n < F; ++n) {
      const double s_new = (double)n / ((double)F - 1);
      while (seg_end < s_new && k + 1 < F) {
        ++k;
        seg_start += h(k - 1);
        seg_end += h(k);
      }
      s.cplan_k[n] = k;
      s.cplan_w[n] = (s_new - seg_start) / h(k);
    }
  }
}

// generateNewFramesLinear (interpolation.cu:190-227) over all planes of a
// string: image n of the output is w x[k] + (1 - w) x[k-1].
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_linear(AuxArgs a) {
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  const TileCtx tc = tile_of(s, it);
  const int k = s.plan_k[tc.n];
  const double w = s.plan_w[tc.n];
  const T* Xk = (const T*)s.x[a.xin] + (long long)k * s.image_stride;
  const T* Xk1 = Xk - s.image_stride;
  T* O = (T*)s.x[a.xout] + (long long)tc.n * s.image_stride;
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    for (int p = 0; p < s.nplanes; ++p) {
      const long long o = p * s.plane_stride + off;
      const double v = fma(w, (double)Xk[o], (1 - w) * (double)Xk1[o]);
      O[o] = (T)v;
    }
  }
}

// Copies the fixed A planes of a NO_SELF_FIELD string from buffer xout to
// buffer xin after a redistribution, so both field buffers hold the same A.
template <class T>
__global__ __launch_bounds__(kBlock) void k_copy_a(AuxArgs a) {
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  if (!s.nsf) return;
  const TileCtx tc = tile_of(s, it);
  const T* S = (const T*)s.x[a.xout] + (long long)tc.n * s.image_stride;
  T* O = (T*)s.x[a.xin] + (long long)tc.n * s.image_stride;
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    O[off] = S[off];
    O[s.plane_stride + off] = S[s.plane_stride + off];
  }
}

// redistributeCubic (interpolation.cu:292-481) in double, one thread per
// site and plane: divided differences and end third differences
// (computeDeltaX/Y), the Thomas solve of the spline system along the image
// axis (in place of cusparseSgtsv2), and the spline evaluation
// (generateNewFramesCubic).  textbook = 0 keeps the reference's
// (w-1)^3-(w-1) weight on sigma_{k-1}; 1 uses (1-w)^3-(1-w).
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_cubic(AuxArgs a, int textbook) {
  const WorkItem it = a.items[xcd_item(blockIdx.x, a.nitems)];
  const DevString& s = a.strings[it.string];
  const TileCtx tc = tile_of(s, it);
  // The solve runs along the image axis: only the image-0 items of the
  // shared work list take part, one thread per site.
  if (tc.n != 0) return;
  const int F = s.F;
  const double* cb = s.cubic;
  const long long cells = s.plane_stride;
  double* Dl = (double*)s.cubic_rhs;  // [F][nplanes][cells]
  for (int r = 0; r < tc.nrows; ++r) {
    const int i = tc.i, j = tc.j0 + r * tc.rstride;
    if (i >= s.N || j >= s.N) continue;
    const long long off = (long long)j * s.pitch + i;
    for (int p = 0; p < s.nplanes; ++p) {
      const T* Xp = (const T*)s.x[a.xin] + p * s.plane_stride + off;
      auto f = [&](int n) { return (double)Xp[(long long)n * s.image_stride]; };
      auto D = [&](int n) -> double& { return Dl[((long long)n * s.nplanes + p) * cells + off]; };
      auto h = [&](int n) { return cb[4 * n + 3]; };
      for (int n = 1; n < F; ++n) D(n) = (f(n) - f(n - 1)) / h(n);
      const double d3s =
          ((D(3) - D(2)) / (h(3) + h(2)) - (D(2) - D(1)) / (h(2) + h(1))) / (h(3) + h(2) + h(1));
      const double d3e = ((D(F - 1) - D(F - 2)) / (h(F - 1) + h(F - 2)) -
                          (D(F - 2) - D(F - 3)) / (h(F - 2) + h(F - 3))) /
                         (h(F - 1) + h(F - 2) + h(F - 3));
      for (int n = 1; n < F - 1; ++n) D(n) = D(n + 1) - D(n);
      D(0) = h(1) * h(1) * d3s;
      D(F - 1) = -(h(F - 1) * h(F - 1)) * d3e;
      // Thomas forward sweep and back substitution.
      D(0) = D(0) / cb[1];
      for (int n = 1; n < F; ++n) D(n) = (D(n) - cb[4 * n + 2] * D(n - 1)) / cb[4 * n + 1];
      for (int n = F - 2; n >= 0; --n) D(n) = D(n) - cb[4 * n + 0] * D(n + 1);
      T* O = (T*)s.x[a.xout] + p * s.plane_stride + off;
      for (int n = 0; n < F; ++n) {
        const int k = s.cplan_k[n];
        const double w = s.cplan_w[n];
        const double wm = textbook ? ((1 - w) * (1 - w) * (1 - w) - (1 - w))
                                   : ((w - 1) * (w - 1) * (w - 1) - (w - 1));
        const double v = w * f(k) + (1 - w) * f(k - 1) +
                         h(k) * h(k) * ((w * w * w - w) * D(k) + wm * D(k - 1));
        O[(long long)n * s.image_stride] = (T)v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// buildNearestNeighbourMap (nlcg.cu:48-103) for the superconducting (low
// byte) and computational (high byte) domains of one string.
__global__ void k_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch,
                         uint16_t* nn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= N) return;
  auto bits = [&](const int* dom) {
    const int idx = j * N + i;
    int m = 0;
    if (dom[idx] == 0) return 0;
    const bool xr = i != N - 1, xl = i != 0, yu = j != N - 1, yd = j != 0;
    if (xr && dom[idx + 1]) m |= 1 << 0;
    if (xr && yu && dom[idx + 1 + N]) m |= 1 << 1;
    if (yu && dom[idx + N]) m |= 1 << 2;
    if (yu && xl && dom[idx - 1 + N]) m |= 1 << 3;
    if (xl && dom[idx - 1]) m |= 1 << 4;
    if (xl && yd && dom[idx - 1 - N]) m |= 1 << 5;
    if (yd && dom[idx - N]) m |= 1 << 6;
    if (yd && xr && dom[idx - N + 1]) m |= 1 << 7;
    return m;
  };
  nn[(long long)j * pitch + i] = (uint16_t)(bits(sc_dom) | (bits(comp_dom) << 8));
}

// Per-image scalar maintenance (graph-safe, no memset nodes).
__global__ void k_reset_norms(AuxArgs a, int nstrings, int interior_only) {
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  for (int n = threadIdx.x; n < s.F; n += blockDim.x) {
    if (interior_only && (n == 0 || n == s.F - 1)) continue;
    a.sc[s.img0 + n].g_norm2[0] = INFINITY;
    a.sc[s.img0 + n].g_norm2[1] = INFINITY;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill(T* p, long long n, T v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride) p[e] = v;
}

// ---------------------------------------------------------------------------
// Host launchers.
static inline unsigned grid_of(int nitems) { return (unsigned)nitems; }

template <class T>
static hipError_t launch_step_impl(const StepArgs& a, int which, hipStream_t st) {
  if (a.nitems <= 0) return hipSuccess;
  switch (which) {
    case kKernelGrad:
      hipLaunchKernelGGL(k_grad<T>, dim3(grid_of(a.nitems)), dim3(kBlock), 0, st, a);
      break;
    case kKernelPoly:
      hipLaunchKernelGGL(k_poly<T>, dim3(grid_of(a.nitems)), dim3(kBlock), 0, st, a);
      break;
    case kKernelUpdate:
      hipLaunchKernelGGL(k_update<T>, dim3(grid_of(a.nitems)), dim3(kBlock), 0, st, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_step(const StepArgs& a, int which, bool fp32, hipStream_t st) {
  return fp32 ? launch_step_impl<float>(a, which, st) : launch_step_impl<double>(a, which, st);
}

template <class T>
static hipError_t launch_aux_impl(const AuxArgs& a, int which, int arg, hipStream_t st) {
  const dim3 g(grid_of(a.nitems)), b(kBlock);
  switch (which) {
    case kKernelEnergy:
      if (a.nitems > 0) hipLaunchKernelGGL(k_energy<T>, g, b, 0, st, a);
      break;
    case kKernelJB:
      if (a.nitems > 0) hipLaunchKernelGGL(k_jb<T>, g, b, 0, st, a);
      break;
    case kKernelArclen:
      if (a.nitems > 0) hipLaunchKernelGGL(k_arclen<T>, g, b, 0, st, a);
      break;
    case kKernelArcFinalize:
      hipLaunchKernelGGL(k_arc_finalize, dim3((arg + 63) / 64), dim3(64), 0, st, a, arg);
      break;
    case kKernelLinear:
      if (a.nitems > 0) hipLaunchKernelGGL(k_redistribute_linear<T>, g, b, 0, st, a);
      break;
    case kKernelCubic:
      if (a.nitems > 0) hipLaunchKernelGGL(k_redistribute_cubic<T>, g, b, 0, st, a, arg);
      break;
    case kKernelCopyA:
      if (a.nitems > 0) hipLaunchKernelGGL(k_copy_a<T>, g, b, 0, st, a);
      break;
    case kKernelResetNorms:
      hipLaunchKernelGGL(k_reset_norms, dim3(a.nstrings), dim3(64), 0, st, a, a.nstrings, arg);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_aux(const AuxArgs& a, int which, int arg, bool fp32, hipStream_t st) {
  return fp32 ? launch_aux_impl<float>(a, which, arg, st)
              : launch_aux_impl<double>(a, which, arg, st);
}

hipError_t launch_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch, uint16_t* nn,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_nn_map, dim3((N + 63) / 64, N), dim3(64), 0, st, sc_dom, comp_dom, N,
                     pitch, nn);
  return hipGetLastError();
}

hipError_t launch_fill(void* p, long long n, bool fp32, double v, hipStream_t st) {
  const unsigned blocks = (unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 4096);
  if (n <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(kBlock), 0, st, (float*)p, n, (float)v);
  else
    hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(kBlock), 0, st, (double*)p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_fill<unsigned>, dim3(blocks), dim3(kBlock), 0, st, p, n, 0u);
  return hipGetLastError();
}

}  // namespace mepgl
