// HIP kernels of the MI355X gauged-string-method relaxation loop (gfx950).
//
// Work decomposition: one wavefront per work item = one 64-column strip of
// one row segment of one image of one string.  Lane l owns column
// 64*strip + l; the wavefront marches along the rows of its segment (down in
// even segments, up in odd ones), keeping the rows j-1, j, j+1 it needs in a
// register window and prefetching the next-but-one row while it computes
// row j.  Neighbours along x come from the adjacent lanes through DPP wave
// shifts; the two columns just outside the strip are loaded by lanes 0 and
// 63 only.  Every field value is therefore read from HBM once per launch,
// plus the halo rows and columns the strip shares with its neighbours,
// which L2 serves: adjacent work items run at the same time on the same
// XCD, and neighbouring segments reach their shared rows together.
//
// One NLCG step (nlcgSteps, nlcg.cu:1478-1583) is two launches:
//   G  fused pending update x += alpha d (updateFields), gradient
//      (gradientPsi + gradientA), |g|^2 and g.(g - g_old) partials
//      (gradientDensity).
//   P  direction d = -g + beta d (computeStepDirection), evaluated on the
//      stencil, line-search quartic partials (polynomialExpansion).
// Each is followed by a finishing launch with one workgroup per image that
// reduces the partials in a fixed order: beta after G; A..E and the line
// search (lineSearch, a parallel scan over a host-built table) after P.
// Reductions are deterministic (fixed partial order, fixed trees), with
// two-sum compensation in the mixed-precision path.
#include <float.h>

#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "stencil.hpp"

namespace mepgl {

// ---------------------------------------------------------------------------
// Work-item mapping: consecutive items go to the same XCD so that the halo
// rows and columns a strip shares with its neighbours are served by that
// XCD's L2.
__device__ __forceinline__ int xcd_item(int b, int nb) {
  const int q = nb / kNumXcd, r = nb % kNumXcd;
  const int k = b % kNumXcd, l = b / kNumXcd;
  return k * q + (k < r ? k : r) + l;
}

// The work item of this workgroup, made wave-uniform so that the descriptor
// behind it is read with scalar loads.
__device__ __forceinline__ WorkItem load_item(const WorkItem* items, int k) {
  const WorkItem w = items[k];
  WorkItem u;
  u.string = __builtin_amdgcn_readfirstlane(w.string);
  u.image = __builtin_amdgcn_readfirstlane(w.image);
  u.tile = __builtin_amdgcn_readfirstlane(w.tile);
  u.pad = 0;
  return u;
}

// Pointers read from descriptors are generic; the hot kernels cast them to
// the global address space so that loads and stores compile to global_*
// instructions instead of flat_*.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

// Two-sum accumulation (Knuth): exact rounding error kept in c.
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

// Wavefront reduction of K (sum, compensation) pairs; every lane ends with
// the same result (butterfly, operands combined in lane order).
template <int K, bool COMP>
__device__ __forceinline__ void wave_reduce(double (&s)[K], double (&c)[K]) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const bool lo = (threadIdx.x & off) == 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double so = __shfl_xor(s[k], off);
      if (COMP) {
        const double co = __shfl_xor(c[k], off);
        double x = lo ? s[k] : so, y = lo ? so : s[k];
        double cx = lo ? c[k] : co, cy = lo ? co : c[k];
        two_sum_acc(x, cx, y);
        s[k] = x;
        c[k] = cx + cy;
      } else {
        s[k] = lo ? s[k] + so : so + s[k];
      }
    }
  }
}

// Publishes this wavefront's K (sum, compensation) pairs as partial `tile`
// of image n.  Partials are combined by a separate finishing launch: a
// cross-wavefront handshake inside the kernel would need an agent-scope
// release per wavefront, which on a multi-XCD device writes back the L2.
template <int K, bool COMP>
__device__ __forceinline__ void publish_partial(double (&s)[K], double (&c)[K], double* part,
                                                int n, int tile, int ntiles) {
  wave_reduce<K, COMP>(s, c);
  if (threadIdx.x == 0) {
    double* mine = part + ((long long)n * ntiles + tile) * (2 * K);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      mine[k] = s[k];
      mine[K + k] = c[k];
    }
  }
}

// Fixed-order total of the ntiles partials of image n (every lane).
template <int K, bool COMP>
__device__ __forceinline__ void reduce_partials(const double* part, int n, int ntiles,
                                                double (&fin)[K]) {
  double ss[K], cc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) ss[k] = cc[k] = 0.0;
  const double* base = part + (long long)n * ntiles * (2 * K);
  // Each lane takes tiles lane, lane + 64, ... in that order; the loads of
  // four of its tiles are issued before any is accumulated, so a lane waits
  // for memory once per four tiles instead of once per tile.
  constexpr int U = 4;
  for (int t0 = threadIdx.x; t0 < ntiles; t0 += U * kWave) {
    double v[U][K], w[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = t0 + u * kWave;
#pragma unroll
      for (int k = 0; k < K; ++k) v[u][k] = w[u][k] = 0.0;
      if (t < ntiles) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          v[u][k] = base[(long long)t * 2 * K + k];
          if (COMP) w[u][k] = base[(long long)t * 2 * K + K + k];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (t0 + u * kWave < ntiles) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          acc<COMP>(ss[k], cc[k], v[u][k]);
          if (COMP) cc[k] += w[u][k];
        }
      }
    }
  }
  wave_reduce<K, COMP>(ss, cc);
#pragma unroll
  for (int k = 0; k < K; ++k) fin[k] = ss[k] + cc[k];
}

// ---------------------------------------------------------------------------
// Lane shifts across the wavefront (DPP wave_shr:1 / wave_shl:1).  Lanes
// without a source keep `edge`, which holds the column outside the strip.
__device__ __forceinline__ float from_left(float v, float edge) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v, float edge) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_left(double v, double edge) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(edge), __double2loint(v), 0x138, 0xF,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), __double2hiint(v), 0x138, 0xF,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v, double edge) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(edge), __double2loint(v), 0x130, 0xF,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), __double2hiint(v), 0x130, 0xF,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}

// Strip geometry of a work item.
struct Strip {
  int n, tile;
  int lane, i, ic;   // lane, column, column clamped into the grid
  int hc;            // halo column of lanes 0 and 63 (own column elsewhere)
  bool edge;         // lane 0 or 63
  int j0, j1;        // rows [j0, j1) of the segment
};

__device__ __forceinline__ Strip strip_of(const DevString& s, const WorkItem& it) {
  Strip t;
  t.n = it.image;
  t.tile = it.tile;
  const int sx = it.tile % s.tiles_x, sy = it.tile / s.tiles_x;
  t.lane = threadIdx.x;
  const int x0 = sx * kTileX;
  t.i = x0 + t.lane;
  t.ic = min(t.i, s.N - 1);
  t.edge = t.lane == 0 || t.lane == kWave - 1;
  t.hc = t.lane == 0 ? max(x0 - 1, 0) : (t.lane == kWave - 1 ? min(x0 + kWave, s.N - 1) : t.ic);
  t.j0 = sy * s.seg_rows;
  t.j1 = min(t.j0 + s.seg_rows, s.N);
  return t;
}

// Segments alternate their march direction: even segments march down, odd
// ones up.  Two segments that share a boundary then read the rows around it
// at about the same time (both at their start or both at their end), and
// vertically adjacent segments run concurrently on one XCD (xcd_item), so
// the second read of each shared row is served by that XCD's L2.
__device__ __forceinline__ bool marches_up(const DevString& s, const WorkItem& it) {
  return __builtin_amdgcn_readfirstlane((it.tile / s.tiles_x) & 1) != 0;
}

__device__ __forceinline__ int clampr(int j, int N) { return j < 0 ? 0 : (j >= N ? N - 1 : j); }

__device__ __forceinline__ double hval(const DevString& s, long long off) {
  return s.h.p ? s.h.p[off] : s.h.v;
}

// ---------------------------------------------------------------------------
// Buffer access for the G and P kernels.  Each field set of one image is one
// buffer resource (base in scalar registers); an element is addressed by a
// 32-bit byte offset of (row, column) in a VGPR plus the plane's byte
// offset in an SGPR, so no 64-bit address arithmetic runs per load.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // raw untyped buffer (CDNA)

__device__ __forceinline__ Rsrc rsrc_of(const void* base, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           (int)(bytes > 0x7fffffffLL ? 0x7fffffff : bytes),
                                           kRsrcWord3);
}

__device__ __forceinline__ double bld(double, Rsrc r, unsigned vo, unsigned so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ float bld(float, Rsrc r, unsigned vo, unsigned so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
template <class T>
__device__ __forceinline__ T bload(Rsrc r, unsigned vo, unsigned so) {
  return bld(T(), r, vo, so);
}
// Stores are non-temporal (nt): the next reader is another launch, after
// far more data than the caches hold has streamed through.
constexpr int kStoreNT = 2;
__device__ __forceinline__ void bstore(double v, Rsrc r, unsigned vo, unsigned so) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vo, so, kStoreNT);
}
__device__ __forceinline__ void bstore(float v, Rsrc r, unsigned vo, unsigned so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, kStoreNT);
}
__device__ __forceinline__ uint16_t bload_u16(Rsrc r, unsigned vo) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, vo, 0, 0);
}

// Per-site parameters of one row when some parameter is a plane: a, b,
// m_xx, m_yy per component and the applied field at the five stencil
// points (centre, +x, -x, +y, -y).
template <class T, int NC, bool PL>
struct RowParams {
  __device__ __forceinline__ void load(const ParamView<NC>&, gptr<const double>, long long, int,
                                       int, int, const Strip&) {}
};
template <class T, int NC>
struct RowParams<T, NC, true> {
  double prm[NC][4];
  T h[5];
  __device__ __forceinline__ void load(const ParamView<NC>& pv, gptr<const double> hp, long long off,
                                       int j, int N, int P_, const Strip& t) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      prm[c][0] = param_at<T>(pv.a[c], off);
      prm[c][1] = param_at<T>(pv.b[c], off);
      prm[c][2] = param_at<T>(pv.mxx[c], off);
      prm[c][3] = param_at<T>(pv.myy[c], off);
    }
    if (hp) {
      const int ip = min(t.ic + 1, N - 1), im = max(t.ic - 1, 0);
      h[0] = (T)hp[off];
      h[1] = (T)hp[(long long)j * P_ + ip];
      h[2] = (T)hp[(long long)j * P_ + im];
      h[3] = (T)hp[(long long)clampr(j + 1, N) * P_ + t.ic];
      h[4] = (T)hp[(long long)clampr(j - 1, N) * P_ + t.ic];
    }
  }
};

template <class T, int NC>
__device__ __forceinline__ void comp_from_params(const double (&prm)[NC][4], const double (&q)[NC],
                                                 CompConst<T> (&k)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double a = prm[c][0], b = prm[c][1], mx = prm[c][2], my = prm[c][3];
    k[c].q = (T)q[c];
    k[c].two_a = (T)(2.0 * a);
    k[c].two_b = (T)(2.0 * b);
    k[c].b = (T)b;
    k[c].ab = (T)(a / b);
    k[c].hx = (T)(0.5 / mx);
    k[c].hy = (T)(0.5 / my);
  }
}

// Byte offsets of a strip row: own column and halo column.
struct RowOff {
  unsigned own, halo;
};
template <class T>
__device__ __forceinline__ RowOff row_off(int row, int P_, const Strip& t) {
  return {(unsigned)((row * P_ + t.ic) * (int)sizeof(T)), (unsigned)((row * P_ + t.hc) * (int)sizeof(T))};
}

// Makes a value live in a register at this point of the program: the
// compiler may not defer its computation (and the wait for the loads it
// reads) past the point.
template <class T>
__device__ __forceinline__ void pin(T& v) {
  asm volatile("" : "+v"(v));
}
// A compiler barrier for memory operations: later stores are not moved above
// it (pinned computations before it stay before the stores).
__device__ __forceinline__ void memory_fence_compiler() { asm volatile("" ::: "memory"); }

// ---------------------------------------------------------------------------
// G kernel.
template <class T, int NC>
struct GRaw {
  static constexpr int P = 2 + 2 * NC;
  T x[P], d[P], hx[P], hd[P];
  uint16_t nn;
};

template <class T, int NC>
struct GRow {
  T ax, ay, axl, axr, ayl, ayr;  // A and its x-neighbours
  T u[NC], v[NC];                // psi
  T eu[NC], ev[NC];              // psi outside the strip (lanes 0 / 63)
  uint16_t nn;
};

// Buffers of one G work item.
struct GBufs {
  Rsrc x, xo, d, g, go, nn;
  unsigned ps;  // plane stride in bytes
};

template <class T, int NC, bool UPD, bool NSF>
__device__ __forceinline__ void g_load(GRaw<T, NC>& r, const GBufs& b, int row, int P_,
                                       const Strip& t, unsigned nnoff) {
  const RowOff o = row_off<T>(row, P_, t);
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) {
    r.x[p] = bload<T>(b.x, o.own, p * b.ps);
    r.d[p] = (UPD && !(NSF && p < 2)) ? bload<T>(b.d, o.own, p * b.ps) : (T)0;
  }
  // Halo columns: only lanes 0 and 63 read the columns outside the strip
  // (exec-masked loads); the other lanes' halo values are never used.
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) r.hx[p] = r.hd[p] = (T)0;
  if (t.edge) {
#pragma unroll
    for (int p = 0; p < GRaw<T, NC>::P; ++p) {
      r.hx[p] = bload<T>(b.x, o.halo, p * b.ps);
      r.hd[p] = (UPD && !(NSF && p < 2)) ? bload<T>(b.d, o.halo, p * b.ps) : (T)0;
    }
  }
  r.nn = bload_u16(b.nn, nnoff);
}

template <class T, int NC>
__device__ __forceinline__ void pin_row(GRow<T, NC>& w) {
  pin(w.ax), pin(w.ay), pin(w.axl), pin(w.axr), pin(w.ayl), pin(w.ayr);
#pragma unroll
  for (int q = 0; q < NC; ++q) pin(w.u[q]), pin(w.v[q]), pin(w.eu[q]), pin(w.ev[q]);
  pin(w.nn);
}

template <class T, int NC>
__device__ __forceinline__ void g_process(GRow<T, NC>& w, const GRaw<T, NC>& r, T al, bool upd) {
  T c[GRaw<T, NC>::P], e[GRaw<T, NC>::P];
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) {
    c[p] = upd ? fma(al, r.d[p], r.x[p]) : r.x[p];
    e[p] = upd ? fma(al, r.hd[p], r.hx[p]) : r.hx[p];
  }
  w.ax = c[PAX];
  w.ay = c[PAY];
  w.axl = from_left(c[PAX], e[PAX]);
  w.axr = from_right(c[PAX], e[PAX]);
  w.ayl = from_left(c[PAY], e[PAY]);
  w.ayr = from_right(c[PAY], e[PAY]);
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.u[q] = c[PU1 + 2 * q];
    w.v[q] = c[PV1 + 2 * q];
    w.eu[q] = e[PU1 + 2 * q];
    w.ev[q] = e[PV1 + 2 * q];
  }
  w.nn = r.nn;
}

// Own-site values of a row that only the row's own computation needs: the
// old gradient (Polak-Ribiere), the frozen psi gradient (Maxwell mode) and,
// with parameter planes, the row's parameters.
template <class T, int NC, bool PL>
struct GSide {
  static constexpr int P = 2 + 2 * NC;
  T go[P];
  T gst[NC][2];
  RowParams<T, NC, PL> rp;
};

// The old gradient is read unconditionally (the diagnostic entry point that
// does not need it still reads a valid buffer), so the batch has no
// data-dependent branches.
template <class T, int NC, bool NSF, bool PL, bool MW>
__device__ __forceinline__ void g_side_load(GSide<T, NC, PL>& sd, const GBufs& b, int row, int P_,
                                            const Strip& t, const ParamView<NC>& pv,
                                            gptr<const double> hp, int N) {
  const unsigned o = (unsigned)((row * P_ + t.ic) * (int)sizeof(T));
#pragma unroll
  for (int p = 0; p < GSide<T, NC, PL>::P; ++p)
    sd.go[p] = !(NSF && p < 2) ? bload<T>(b.go, o, p * b.ps) : (T)0;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    sd.gst[q][0] = MW ? bload<T>(b.g, o, (PU1 + 2 * q) * b.ps) : (T)0;
    sd.gst[q][1] = MW ? bload<T>(b.g, o, (PV1 + 2 * q) * b.ps) : (T)0;
  }
  sd.rp.load(pv, hp, (long long)row * P_ + t.ic, row, N, P_, t);
}

template <class T, int NC, bool PL>
__device__ __forceinline__ void pin_side(GSide<T, NC, PL>& sd) {
#pragma unroll
  for (int p = 0; p < GSide<T, NC, PL>::P; ++p) pin(sd.go[p]);
#pragma unroll
  for (int q = 0; q < NC; ++q) pin(sd.gst[q][0]), pin(sd.gst[q][1]);
  if constexpr (PL) {
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) pin(sd.rp.prm[q][k]);
#pragma unroll
    for (int k = 0; k < 5; ++k) pin(sd.rp.h[k]);
  }
}

template <class T, int NC, bool UPD, bool NSF, bool PL, bool MW>
__device__ __forceinline__ void grad_body(const StepArgs& a, const DevString& s, const WorkItem& it) {
  constexpr bool COMP = sizeof(T) == 4;
  const Strip t = strip_of(s, it);
  const int N = s.N, P_ = s.pitch;
  const long long ibytes = s.image_stride * (long long)sizeof(T);
  const long long ib = (long long)t.n * s.image_stride;
  GBufs b;
  b.x = rsrc_of((const T*)s.x[a.xin] + ib, ibytes);
  b.xo = rsrc_of((const T*)s.x[a.xout] + ib, ibytes);
  b.d = rsrc_of((const T*)s.d[a.din] + ib, ibytes);
  b.g = rsrc_of((const T*)(a.scratch ? s.gs : s.g[a.gcur]) + ib, ibytes);
  b.go = rsrc_of((const T*)s.g[a.gold] + ib, ibytes);
  b.nn = rsrc_of(s.nn, (long long)s.N * s.pitch * (long long)sizeof(uint16_t));
  b.ps = (unsigned)(s.plane_stride * (long long)sizeof(T));
  const ImageScalars& sc = a.sc[s.img0 + t.n];
  const T al = UPD ? (T)sc.alpha : (T)0;
  constexpr bool do_a = !NSF;
  const bool want_go = !a.scratch;
  const bool store = t.i < N;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  CompConst<T> kconst[NC];
  if (!PL) load_comp_const<T, NC>(s, 0, kconst);
  const ParamView<NC> pv = param_view<NC>(s);
  gptr<const double> __restrict__ hp = as_global(s.h.p);
  const T hv = (T)s.h.v;
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  auto nnoff = [&](int row) { return (unsigned)((row * P_ + t.ic) * (int)sizeof(uint16_t)); };

  // The march is the only part that depends on the direction
  // (marches_up); everything it needs from the string's descriptor is
  // read above, before the branch and before any store, so it stays in
  // scalar registers.
  const bool up = marches_up(s, it);
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  auto march = [&](auto up_tag) {
    constexpr bool UP = decltype(up_tag)::value;
    constexpr int D = UP ? -1 : 1;  // march direction in rows
    // The window is three row slots used in rotation: (previous, current,
    // next) = (w[IP], w[IC], w[IN]) in the order of the march, with the side
    // values of each row in the slot of the same index.  Each row writes the
    // newly loaded row into the previous row's slot once the current row is
    // computed, and the next row uses the slots shifted by one, so no register
    // copies advance the window.  A segment marches down (rows j0 .. j1-1) or,
    // when UP, up (rows j1-1 .. j0); the previous row of the march is then
    // row j + 1.
    const int jb = UP ? t.j1 - 1 : t.j0;  // first row of the march
    GRaw<T, NC> raw;
    GRow<T, NC> w[3];
    GSide<T, NC, PL> sd[3];
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb - D, N), P_, t, nnoff(clampr(jb - D, N)));
    g_process<T, NC>(w[0], raw, al, UPD);
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb, N), P_, t, nnoff(clampr(jb, N)));
    g_process<T, NC>(w[1], raw, al, UPD);
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb + D, N), P_, t, nnoff(clampr(jb + D, N)));
    g_process<T, NC>(w[2], raw, al, UPD);
    g_side_load<T, NC, NSF, PL, MW>(sd[1], b, jb, P_, t, pv, hp, N);

    // Row j: issue one batch of loads (the fields of row j + 2D, the own-site
    // values of row j + D), compute row j from registers, turn the batch into
    // window form, then store row j.  Every load is consumed before this
    // row's stores are issued, and the last row (no batch) is a separate copy
    // of the body, so the loads in flight at each point are known exactly.
    // Rows past the end of the segment (padding of the last rotation) are
    // computed but neither stored nor summed.
    auto row = [&](int j, auto ip_tag, auto ic_tag, auto in_tag, auto batch_tag) {
      constexpr int IP = decltype(ip_tag)::value, IC = decltype(ic_tag)::value;
      constexpr int IN = decltype(in_tag)::value;
      constexpr bool BATCH = decltype(batch_tag)::value;
      GRow<T, NC>& wm = w[UP ? IN : IP];  // row j - 1
      GRow<T, NC>& wc = w[IC];
      GRow<T, NC>& wpl = w[UP ? IP : IN];  // row j + 1
      const bool live = store && (UP ? j >= t.j0 : j < t.j1);
      if constexpr (BATCH) {
        g_load<T, NC, UPD, NSF>(raw, b, clampr(j + 2 * D, N), P_, t,
                                nnoff(clampr(j + 2 * D, N)));
        g_side_load<T, NC, NSF, PL, MW>(sd[IN], b, clampr(j + D, N), P_, t, pv, hp, N);
      }
      __builtin_amdgcn_sched_barrier(0);

      Nbhd<T, NC> nb;
      nb.ax[0][0] = wm.axl; nb.ax[0][1] = wm.ax; nb.ax[0][2] = wm.axr;
      nb.ax[1][0] = wc.axl; nb.ax[1][1] = wc.ax; nb.ax[1][2] = wc.axr;
      nb.ax[2][0] = wpl.axl; nb.ax[2][1] = wpl.ax; nb.ax[2][2] = wpl.axr;
      nb.ay[0][0] = wm.ayl; nb.ay[0][1] = wm.ay; nb.ay[0][2] = wm.ayr;
      nb.ay[1][0] = wc.ayl; nb.ay[1][1] = wc.ay; nb.ay[1][2] = wc.ayr;
      nb.ay[2][0] = wpl.ayl; nb.ay[2][1] = wpl.ay; nb.ay[2][2] = wpl.ayr;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        nb.u[q][0] = wc.u[q];
        nb.v[q][0] = wc.v[q];
        nb.u[q][1] = from_right(wc.u[q], wc.eu[q]);
        nb.v[q][1] = from_right(wc.v[q], wc.ev[q]);
        nb.u[q][2] = from_left(wc.u[q], wc.eu[q]);
        nb.v[q][2] = from_left(wc.v[q], wc.ev[q]);
        nb.u[q][3] = wpl.u[q];
        nb.v[q][3] = wpl.v[q];
        nb.u[q][4] = wm.u[q];
        nb.v[q][4] = wm.v[q];
      }
      const GSide<T, NC, PL>& cur = sd[IC];
      CompConst<T> k[NC];
      if constexpr (PL) {
        comp_from_params<T, NC>(cur.rp.prm, pv.q, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) nb.h[q] = hp ? cur.rp.h[q] : hv;
      } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) k[q] = kconst[q];
#pragma unroll
        for (int q = 0; q < 5; ++q) nb.h[q] = hv;
      }

      T gp[NC][2], ga[2] = {0, 0};
      site_gradient<T, NC>(nb, k, eta, gam, del, idx, dx2, wc.nn & 0xff, wc.nn >> 8, !MW, do_a,
                           gp, ga);
      if constexpr (MW) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          gp[q][0] = cur.gst[q][0];
          gp[q][1] = cur.gst[q][1];
        }
      }
      if (live && want_go) {
        // gradientDensity, nlcg.cu:919-949
        double n2 = 0, dg = 0;
        if (do_a) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const double gv = (double)ga[q], go_ = (double)cur.go[q];
            n2 += gv * gv;
            dg += gv * (gv - go_);
          }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const double gv = (double)gp[q][r], go_ = (double)cur.go[PU1 + 2 * q + r];
            n2 += gv * gv;
            dg += gv * (gv - go_);
          }
        }
        // Per-lane running sums over the segment's rows in double; the
        // compensation (mixed) starts at the cross-lane and cross-tile sums.
        acc_s[0] += n2;
        acc_s[1] += dg;
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (BATCH) {
        g_process<T, NC>(w[IP], raw, al, UPD);
        pin_row<T, NC>(w[IP]);
        pin_side<T, NC, PL>(sd[IN]);
        memory_fence_compiler();
      }
      __builtin_amdgcn_sched_barrier(0);
      if (live) {
        const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
        if constexpr (!MW) {
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            bstore(gp[q][0], b.g, o, (PU1 + 2 * q) * b.ps);
            bstore(gp[q][1], b.g, o, (PV1 + 2 * q) * b.ps);
          }
        }
        if (do_a) {
          bstore(ga[0], b.g, o, PAX * b.ps);
          bstore(ga[1], b.g, o, PAY * b.ps);
        }
        if (UPD) {
          if (do_a) {
            bstore(wc.ax, b.xo, o, PAX * b.ps);
            bstore(wc.ay, b.xo, o, PAY * b.ps);
          }
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            bstore(wc.u[q], b.xo, o, (PU1 + 2 * q) * b.ps);
            bstore(wc.v[q], b.xo, o, (PV1 + 2 * q) * b.ps);
          }
        }
      }
    };
    // All rows of the segment but the last carry a batch; they run in
    // rotations of three, the last rotation padded past the segment's end when
    // their count is not a multiple of three.
    const int groups = (t.j1 - 1 - t.j0 + 2) / 3;
    int j = jb;
#pragma clang loop unroll(disable)
    for (int g = 0; g < groups; ++g, j += 3 * D) {
      row(j, S0(), S1(), S2(), std::true_type());
      row(j + D, S1(), S2(), S0(), std::true_type());
      row(j + 2 * D, S2(), S0(), S1(), std::true_type());
    }
    row(j, S0(), S1(), S2(), std::false_type());
  };
  if (up)
    march(std::true_type());
  else
    march(std::false_type());
  if (a.scratch) return;

  publish_partial<2, COMP>(acc_s, acc_c, s.part_g, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T, int NC, bool UPD, bool NSF, bool PL, bool MW>
__global__ __launch_bounds__(kBlock) void k_grad(StepArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  grad_body<T, NC, UPD, NSF, PL, MW>(a, a.strings[it.string], it);
}

// ---------------------------------------------------------------------------
// P kernel.
template <class T, int NC>
struct PRaw {
  static constexpr int P = 2 + 2 * NC;
  T x[P], g[P], d[P], hx[P], hg[P], hd[P];
  uint16_t nn;
};

template <class T, int NC>
struct PRow {
  T ax, ay, dax, day;              // A and its new direction
  T eax, eay, edax, eday;          // the same outside the strip (lanes 0 / 63)
  T u[NC], v[NC], du[NC], dv[NC];  // psi and its new direction
  T eu[NC], ev[NC], edu[NC], edv[NC];
  uint16_t nn;
};

struct PBufs {
  Rsrc x, g, d, dout, nn;
  unsigned ps;
};

template <class T, int NC, bool NSF>
__device__ __forceinline__ void p_load(PRaw<T, NC>& r, const PBufs& b, int row, int P_,
                                       const Strip& t, unsigned nnoff) {
  const RowOff o = row_off<T>(row, P_, t);
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) {
    const bool dir = !(NSF && p < 2);
    r.x[p] = bload<T>(b.x, o.own, p * b.ps);
    r.g[p] = dir ? bload<T>(b.g, o.own, p * b.ps) : (T)0;
    r.d[p] = dir ? bload<T>(b.d, o.own, p * b.ps) : (T)0;
  }
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) r.hx[p] = r.hg[p] = r.hd[p] = (T)0;
  if (t.edge) {
#pragma unroll
    for (int p = 0; p < PRaw<T, NC>::P; ++p) {
      const bool dir = !(NSF && p < 2);
      r.hx[p] = bload<T>(b.x, o.halo, p * b.ps);
      r.hg[p] = dir ? bload<T>(b.g, o.halo, p * b.ps) : (T)0;
      r.hd[p] = dir ? bload<T>(b.d, o.halo, p * b.ps) : (T)0;
    }
  }
  r.nn = bload_u16(b.nn, nnoff);
}

// d = -g + beta d  (computeStepDirection, nlcg.cu:985-993)
template <class T, int NC>
__device__ __forceinline__ void pin_row(PRow<T, NC>& w) {
  pin(w.ax), pin(w.ay), pin(w.dax), pin(w.day), pin(w.eax), pin(w.eay), pin(w.edax), pin(w.eday);
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    pin(w.u[q]), pin(w.v[q]), pin(w.du[q]), pin(w.dv[q]);
    pin(w.eu[q]), pin(w.ev[q]), pin(w.edu[q]), pin(w.edv[q]);
  }
  pin(w.nn);
}

template <class T, int NC>
__device__ __forceinline__ void p_process(PRow<T, NC>& w, const PRaw<T, NC>& r, T beta) {
  T dn[PRaw<T, NC>::P], en[PRaw<T, NC>::P];
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) {
    dn[p] = fma(beta, r.d[p], -r.g[p]);
    en[p] = fma(beta, r.hd[p], -r.hg[p]);
  }
  w.ax = r.x[PAX];
  w.ay = r.x[PAY];
  w.eax = r.hx[PAX];
  w.eay = r.hx[PAY];
  w.dax = dn[PAX];
  w.day = dn[PAY];
  w.edax = en[PAX];
  w.eday = en[PAY];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.u[q] = r.x[PU1 + 2 * q];
    w.v[q] = r.x[PV1 + 2 * q];
    w.eu[q] = r.hx[PU1 + 2 * q];
    w.ev[q] = r.hx[PV1 + 2 * q];
    w.du[q] = dn[PU1 + 2 * q];
    w.dv[q] = dn[PV1 + 2 * q];
    w.edu[q] = en[PU1 + 2 * q];
    w.edv[q] = en[PV1 + 2 * q];
  }
  w.nn = r.nn;
}

template <class T, int NC, bool NSF, bool PL>
__device__ __forceinline__ void poly_body(const StepArgs& a, const DevString& s, const WorkItem& it) {
  constexpr bool COMP = sizeof(T) == 4;
  const Strip t = strip_of(s, it);
  const int N = s.N, P_ = s.pitch;
  const long long ibytes = s.image_stride * (long long)sizeof(T);
  const long long ib = (long long)t.n * s.image_stride;
  PBufs b;
  b.x = rsrc_of((const T*)s.x[a.xin] + ib, ibytes);
  b.g = rsrc_of((const T*)s.g[a.gcur] + ib, ibytes);
  b.d = rsrc_of((const T*)s.d[a.din] + ib, ibytes);
  b.dout = rsrc_of((const T*)s.d[a.dout] + ib, ibytes);
  b.nn = rsrc_of(s.nn, (long long)s.N * s.pitch * (long long)sizeof(uint16_t));
  b.ps = (unsigned)(s.plane_stride * (long long)sizeof(T));
  const ImageScalars& sc = a.sc[s.img0 + t.n];
  const T beta = (T)sc.beta;
  constexpr bool nsf = NSF;
  const bool exact = s.poly_exact;
  const bool store = t.i < N;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  CompConst<T> kconst[NC];
  if (!PL) load_comp_const<T, NC>(s, 0, kconst);
  const ParamView<NC> pv = param_view<NC>(s);
  gptr<const double> __restrict__ hp = as_global(s.h.p);
  const T hv = (T)s.h.v;
  double acc_s[5] = {0, 0, 0, 0, 0}, acc_c[5] = {0, 0, 0, 0, 0};
  auto nnoff = [&](int row) { return (unsigned)((row * P_ + t.ic) * (int)sizeof(uint16_t)); };

  // The march is the only part that depends on the direction
  // (marches_up); everything it needs from the string's descriptor is
  // read above, before the branch and before any store, so it stays in
  // scalar registers.
  const bool up = marches_up(s, it);
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  auto march = [&](auto up_tag) {
    constexpr bool UP = decltype(up_tag)::value;
    constexpr int D = UP ? -1 : 1;  // march direction in rows
    // Three row slots in rotation, as in the G kernel: w[IP] is the previous
    // row of the march, w[IC] the current, w[IN] the next; the parameters of a
    // row (parameter planes) sit in the slot of the same index.  Of row j - 1
    // only A_x and its direction are read.
    const int jb = UP ? t.j1 - 1 : t.j0;  // first row of the march
    PRaw<T, NC> raw;
    PRow<T, NC> w[3];
    RowParams<T, NC, PL> rp[3];
    p_load<T, NC, NSF>(raw, b, clampr(jb - D, N), P_, t, nnoff(clampr(jb - D, N)));
    p_process<T, NC>(w[0], raw, beta);
    p_load<T, NC, NSF>(raw, b, clampr(jb, N), P_, t, nnoff(clampr(jb, N)));
    p_process<T, NC>(w[1], raw, beta);
    p_load<T, NC, NSF>(raw, b, clampr(jb + D, N), P_, t, nnoff(clampr(jb + D, N)));
    p_process<T, NC>(w[2], raw, beta);
    rp[1].load(pv, hp, (long long)jb * P_ + t.ic, jb, N, P_, t);

    // Row j: one batch of loads (the fields of row j + 2D, the parameters of
    // row j + D), the quartic of row j from registers, the batch into the
    // previous row's slot, then the stores of row j (see the G loop).
    auto row = [&](int j, auto ip_tag, auto ic_tag, auto in_tag, auto batch_tag) {
      constexpr int IP = decltype(ip_tag)::value, IC = decltype(ic_tag)::value;
      constexpr int IN = decltype(in_tag)::value;
      constexpr bool BATCH = decltype(batch_tag)::value;
      PRow<T, NC>& wm = w[UP ? IN : IP];  // row j - 1
      PRow<T, NC>& wc = w[IC];
      PRow<T, NC>& wpl = w[UP ? IP : IN];  // row j + 1
      const bool live = store && (UP ? j >= t.j0 : j < t.j1);
      if constexpr (BATCH) {
        p_load<T, NC, NSF>(raw, b, clampr(j + 2 * D, N), P_, t, nnoff(clampr(j + 2 * D, N)));
        const int r1 = clampr(j + D, N);
        rp[IN].load(pv, hp, (long long)r1 * P_ + t.ic, r1, N, P_, t);
      }
      __builtin_amdgcn_sched_barrier(0);

      PolyNbhd<T, NC> nb;
      nb.ax[0] = wc.ax;
      nb.ax[1] = from_right(wc.ax, wc.eax);
      nb.ax[2] = 0;
      nb.ax[3] = wpl.ax;
      nb.ax[4] = wm.ax;
      nb.ay[0] = wc.ay;
      nb.ay[1] = from_right(wc.ay, wc.eay);
      nb.ay[2] = from_left(wc.ay, wc.eay);
      nb.ay[3] = wpl.ay;
      nb.ay[4] = 0;
      nb.dax[0] = wc.dax;
      nb.dax[1] = from_right(wc.dax, wc.edax);
      nb.dax[2] = 0;
      nb.dax[3] = wpl.dax;
      nb.dax[4] = wm.dax;
      nb.day[0] = wc.day;
      nb.day[1] = from_right(wc.day, wc.eday);
      nb.day[2] = from_left(wc.day, wc.eday);
      nb.day[3] = wpl.day;
      nb.day[4] = 0;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        nb.u[q][0] = wc.u[q];
        nb.v[q][0] = wc.v[q];
        nb.du[q][0] = wc.du[q];
        nb.dv[q][0] = wc.dv[q];
        nb.u[q][1] = from_right(wc.u[q], wc.eu[q]);
        nb.v[q][1] = from_right(wc.v[q], wc.ev[q]);
        nb.du[q][1] = from_right(wc.du[q], wc.edu[q]);
        nb.dv[q][1] = from_right(wc.dv[q], wc.edv[q]);
        nb.u[q][2] = wpl.u[q];
        nb.v[q][2] = wpl.v[q];
        nb.du[q][2] = wpl.du[q];
        nb.dv[q][2] = wpl.dv[q];
      }
      CompConst<T> k[NC];
      if constexpr (PL) {
        comp_from_params<T, NC>(rp[IC].prm, pv.q, k);
        nb.h0 = hp ? rp[IC].h[0] : hv;
      } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) k[q] = kconst[q];
        nb.h0 = hv;
      }

      if (live) {
        const Quartic<T> qv =
            site_poly<T, NC>(nb, k, eta, gam, del, idx, dx2, wc.nn & 0xff, wc.nn >> 8, nsf, exact);
#pragma unroll
        for (int r = 0; r < 5; ++r) acc_s[r] += (double)qv.c[r];
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (BATCH) {
        p_process<T, NC>(w[IP], raw, beta);
        pin_row<T, NC>(w[IP]);
        if constexpr (PL) {
#pragma unroll
          for (int q = 0; q < NC; ++q)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) pin(rp[IN].prm[q][k2]);
#pragma unroll
          for (int k2 = 0; k2 < 5; ++k2) pin(rp[IN].h[k2]);
        }
        memory_fence_compiler();
      }
      __builtin_amdgcn_sched_barrier(0);
      if (live) {
        const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
        if (!nsf) {
          bstore(wc.dax, b.dout, o, PAX * b.ps);
          bstore(wc.day, b.dout, o, PAY * b.ps);
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          bstore(wc.du[q], b.dout, o, (PU1 + 2 * q) * b.ps);
          bstore(wc.dv[q], b.dout, o, (PV1 + 2 * q) * b.ps);
        }
      }
    };
    const int groups = (t.j1 - 1 - t.j0 + 2) / 3;
    int j = jb;
#pragma clang loop unroll(disable)
    for (int g = 0; g < groups; ++g, j += 3 * D) {
      row(j, S0(), S1(), S2(), std::true_type());
      row(j + D, S1(), S2(), S0(), std::true_type());
      row(j + 2 * D, S2(), S0(), S1(), std::true_type());
    }
    row(j, S0(), S1(), S2(), std::false_type());
  };
  if (up)
    march(std::true_type());
  else
    march(std::false_type());

  publish_partial<5, COMP>(acc_s, acc_c, s.part_p, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T, int NC, bool NSF, bool PL>
__global__ __launch_bounds__(kBlock) void k_poly(StepArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  poly_body<T, NC, NSF, PL>(a, a.strings[it.string], it);
}

// ---------------------------------------------------------------------------
// Finishing launches: one wavefront per image (global image id = block).
__device__ __forceinline__ const DevString& string_of_image(const FinishArgs& f, int& n) {
  const int2 si = f.img[blockIdx.x];
  n = si.y;
  return f.strings[si.x];
}

// After G: |g|^2, g.(g - g_old) and the Polak-Ribiere beta
// (computeStepDirection, nlcg.cu:972-982).
template <bool COMP>
__global__ __launch_bounds__(kBlock) void k_finish_grad(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[2];
  reduce_partials<2, COMP>(s.part_g, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) {
    ImageScalars& sc = f.sc[blockIdx.x];
    sc.g_norm2[f.scur] = fin[0];
    sc.g_delta_g = fin[1];
    double beta = 0;
    if (f.nlcg) {
      beta = fin[1] / sc.g_norm2[f.sold];
      beta = fmax(beta, 0.0);
    }
    sc.beta = beta;
  }
}

// After P: the quartic A..E and the line search (lineSearch).  One
// 1024-thread workgroup per image: wavefront 0 reduces the partials while
// every thread loads its scan candidates p and p + 1024, so the scan takes
// at most two rounds; the first rise is found with an LDS minimum.
constexpr int kFinishPolyBlock = 1024;

template <bool COMP>
__global__ __launch_bounds__(kFinishPolyBlock) void k_finish_poly(FinishArgs f) {
  __shared__ double s_q[5];
  __shared__ int s_first;
  const int2 si = f.img[blockIdx.x];
  const DevString& s = f.strings[si.x];
  const LineSearchTable& t = *f.ls;
  const int tid = threadIdx.x;
  double x[2][2], x4[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int p = min(tid + r * kFinishPolyBlock, kLsSteps - 1);
    x[r][0] = t.x[p];
    x[r][1] = t.x[p + 1];
    x4[r][0] = t.x4[p];
    x4[r][1] = t.x4[p + 1];
  }
  if (tid == 0) s_first = kLsSteps;
  if (tid < kWave) {
    double fin[5];
    reduce_partials<5, COMP>(s.part_p, si.y, s.tiles_x * s.tiles_y, fin);
    if (tid == 0)
#pragma unroll
      for (int k = 0; k < 5; ++k) s_q[k] = fin[k];
  }
  __syncthreads();
  const double A = s_q[0], B = s_q[1], C = s_q[2], D = s_q[3], E = s_q[4];
  // f evaluated without contraction, in the reference's operation order.
  auto fq = [&](double xv, double x4v) {
    double v = __dmul_rn(E, x4v);
    v = __dadd_rn(v, __dmul_rn(__dmul_rn(__dmul_rn(D, xv), xv), xv));
    v = __dadd_rn(v, __dmul_rn(__dmul_rn(C, xv), xv));
    v = __dadd_rn(v, __dmul_rn(B, xv));
    return __dadd_rn(v, A);
  };
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int p = tid + r * kFinishPolyBlock;
    if (p < kLsSteps && fq(x[r][1], x4[r][1]) > fq(x[r][0], x4[r][0])) atomicMin(&s_first, p);
  }
  __syncthreads();
  if (tid == 0) {
    const int first = s_first < kLsSteps ? s_first : kLsSteps - 1;
    ImageScalars& sc = f.sc[blockIdx.x];
    sc.fenergy = A;
    sc.poly[0] = B;
    sc.poly[1] = C;
    sc.poly[2] = D;
    sc.poly[3] = E;
    sc.alpha = t.alpha[first];
    sc.ls_index = first;
  }
}

// After the energy pass (computeFreeEnergy).
__global__ __launch_bounds__(kBlock) void k_finish_energy(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[1];
  reduce_partials<1, true>(s.part_p, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) f.out[blockIdx.x] = fin[0];
}

// After the arclength pass: squared distances to the previous image.
__global__ __launch_bounds__(kBlock) void k_finish_arclen(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[2];
  reduce_partials<2, true>(s.part_a, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].ds2 = fin[0];
    f.sc[blockIdx.x].ds2_jb = fin[1];
  }
}

// ---------------------------------------------------------------------------
// Pending update applied in place (updateFields, nlcg.cu:1372-1396), used
// before anything other than the next gradient reads the fields.
template <class T>
__global__ __launch_bounds__(kBlock) void k_update(StepArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const long long ib = (long long)t.n * s.image_stride;
  T* X = (T*)s.x[a.xin] + ib;
  const T* D = (const T*)s.d[a.din] + ib;
  const T al = (T)a.sc[s.img0 + t.n].alpha;
  const int p0 = s.nsf ? 2 : 0;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
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
__device__ void energy_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  const T* X = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  const double idx = s.inv_dx, dx2 = s.dx2;
  double acc_s[1] = {0}, acc_c[1] = {0};
  if (t.i < s.N) {
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      PolyNbhd<double, NC> nb;
      double um[NC], vm[NC], uym[NC], vym[NC];
      gather5<T, NC>(s, X, t.i, j, nb, um, vm, uym, vym);
      CompConst<double> k[NC];
      load_comp_const<double, NC>(s, off, k);
      const uint16_t nnv = s.nn[off];
      const double e = site_energy<NC>(nb, k, s.eta, s.gamma, s.delta, idx, dx2, nnv & 0xff,
                                       nnv >> 8, s.nsf);
      two_sum_acc(acc_s[0], acc_c[0], e);
    }
  }
  publish_partial<1, true>(acc_s, acc_c, s.part_p, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_energy(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    energy_body<T, 2>(a, s, it);
  else
    energy_body<T, 1>(a, s, it);
}

// computeJB (nlcg.cu:243-380) into planes b, j_1, j_2 of a double buffer
// laid out [image][plane][row][pitch] with 1 + 2*nc planes.
template <class T, int NC>
__device__ void jb_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const T* X = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  double* O = a.jb[it.string] + (long long)t.n * (1 + 2 * NC) * s.plane_stride;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    PolyNbhd<double, NC> nb;
    double um[NC], vm[NC], uym[NC], vym[NC];
    gather5<T, NC>(s, X, t.i, j, nb, um, vm, uym, vym);
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
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    jb_body<T, 2>(a, s, it);
  else
    jb_body<T, 1>(a, s, it);
}

// Arclength partials (arcLengthParametrization + arcLengthParametrizationJB,
// interpolation.cu:39-92): squared distance of image n to image n-1 in the
// plain metric and in the gauge-invariant (b, j) metric, with b and j
// recomputed on the fly for both images.  Image 0 contributes nothing.
template <class T, int NC>
__device__ void arclen_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  if (t.n > 0 && t.i < s.N) {
    const T* X1 = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
    const T* X0 = X1 - s.image_stride;
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      double d2 = 0;
      for (int p = 0; p < 2 + 2 * NC; ++p) {
        const double dv =
            (double)X1[p * s.plane_stride + off] - (double)X0[p * s.plane_stride + off];
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
        gather5<T, NC>(s, m ? X1 : X0, t.i, j, nb, um, vm, uym, vym);
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
  publish_partial<2, true>(acc_s, acc_c, s.part_a, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_arclen(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    arclen_body<T, 2>(a, s, it);
  else
    arclen_body<T, 1>(a, s, it);
}

// computeArcLength (interpolation.cu:126-140) for both metrics, then the
// linear redistribution plan of every new image on s_jb: the segment search
// of generateNewFramesLinear (interpolation.cu:200-219).  One workgroup per
// string.  The square roots run one image per thread; the prefix sums run in
// one thread in the reference's order (so s and s_jb are the reference's
// sums bit for bit); the segment searches run one new image per thread: for
// a non-decreasing s the reference's incremental search stops at the first
// node k >= 1 with s[k] >= s_new, which is what each thread finds directly.
__global__ __launch_bounds__(kArcBlock) void k_arc_finalize(AuxArgs a, int nstrings) {
  __shared__ double ls[kMaxImages], lsj[kMaxImages];
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  const int tid = threadIdx.x;
  if (F < 2 || F > kMaxImages) return;
  for (int n = tid; n < F; n += kArcBlock) {
    ls[n] = n == 0 ? 0.0 : sqrt(sc[n].ds2);
    lsj[n] = n == 0 ? 0.0 : sqrt(sc[n].ds2_jb);
  }
  __syncthreads();
  if (tid == 0) {
    double run = 0, runj = 0;
#pragma unroll 8
    for (int n = 1; n < F; ++n) {
      run += ls[n];
      runj += lsj[n];
      ls[n] = run;
      lsj[n] = runj;
    }
  }
  __syncthreads();
  const double tot = ls[F - 1], totj = lsj[F - 1];
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) {
    ls[n] /= tot;
    lsj[n] /= totj;
    sc[n].s = ls[n];
    sc[n].s_jb = lsj[n];
  }
  __syncthreads();
  // Linear plan (F_new = F).
  for (int n = tid; n < F; n += kArcBlock) {
    const double s_new = (double)n / ((double)F - 1);
    int k = 1;
    while (k + 1 < F && lsj[k] < s_new) ++k;
    const double seg_start = k == 1 ? 0.0 : lsj[k - 1];
    s.plan_k[n] = k;
    s.plan_w[n] = (s_new - seg_start) / (lsj[k] - lsj[k - 1]);
  }
}

// Cubic redistribution plan on the s_jb of the last arclength pass: the
// running sums of the spacings h, the Thomas factors of the spline system
// built by buildHMatrix (interpolation.cu:252-290) and the segment search of
// generateNewFramesCubic (interpolation.cu:373-390).  One workgroup per
// string, launched only before a cubic redistribution; the recurrences run
// in one thread in the reference's order.
__global__ __launch_bounds__(kArcBlock) void k_cubic_plan(AuxArgs a, int nstrings) {
  __shared__ double lsj[kMaxImages], cum[kMaxImages];
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  const ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  const int tid = threadIdx.x;
  if (F < 4 || F > kMaxImages) return;
  double* cb = s.cubic;  // [F][4]: c', m, sub-diagonal, h
  for (int n = tid; n < F; n += kArcBlock) lsj[n] = sc[n].s_jb;
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) cb[4 * n + 3] = n == 0 ? 0.0 : lsj[n] - lsj[n - 1];
  if (tid == 0) {
    auto h = [&](int n) { return n == 0 ? 0.0 : lsj[n] - lsj[n - 1]; };
    cum[0] = 0;
    for (int n = 1; n < F; ++n) cum[n] = cum[n - 1] + h(n);
    double cprev = 0;
    for (int n = 0; n < F; ++n) {
      double dl, d, du;
      if (n == 0) {
        dl = 0; d = -h(1); du = h(1);
      } else if (n == F - 1) {
        dl = h(F - 1); d = -h(F - 1); du = 0;
      } else {
        dl = h(n); d = 2 * (h(n) + h(n + 1)); du = h(n + 1);
      }
      const double m = n == 0 ? d : d - dl * cprev;
      cprev = du / m;
      cb[4 * n + 0] = cprev;
      cb[4 * n + 1] = m;
      cb[4 * n + 2] = dl;
    }
  }
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) {
    const double s_new = (double)n / ((double)F - 1);
    int k = 1;
    while (k + 1 < F && cum[k] < s_new) ++k;
    s.cplan_k[n] = k;
    s.cplan_w[n] = (s_new - cum[k - 1]) / (lsj[k] - lsj[k - 1]);
  }
}

// generateNewFramesLinear (interpolation.cu:190-227) over all planes of a
// string: image n of the output is w x[k] + (1 - w) x[k-1].
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_linear(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const int k = s.plan_k[t.n];
  const double w = s.plan_w[t.n];
  const T* Xk = (const T*)s.x[a.xin] + (long long)k * s.image_stride;
  const T* Xk1 = Xk - s.image_stride;
  T* O = (T*)s.x[a.xout] + (long long)t.n * s.image_stride;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    for (int p = 0; p < s.nplanes; ++p) {
      const long long o = p * s.plane_stride + off;
      O[o] = (T)fma(w, (double)Xk[o], (1 - w) * (double)Xk1[o]);
    }
  }
}

// Copies the fixed A planes of a NO_SELF_FIELD string from buffer xout to
// buffer xin after a redistribution, so both field buffers hold the same A.
template <class T>
__global__ __launch_bounds__(kBlock) void k_copy_a(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (!s.nsf) return;
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const T* S = (const T*)s.x[a.xout] + (long long)t.n * s.image_stride;
  T* O = (T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    O[off] = S[off];
    O[s.plane_stride + off] = S[s.plane_stride + off];
  }
}

// redistributeCubic (interpolation.cu:292-481) in double, one thread per
// site: for each plane, divided differences and end third differences
// (computeDeltaX/Y), the Thomas solve of the spline system along the image
// axis (in place of cusparseSgtsv2), and the spline evaluation
// (generateNewFramesCubic).  textbook = 0 keeps the reference's
// (w-1)^3-(w-1) weight on sigma_{k-1}; 1 uses (1-w)^3-(1-w).
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_cubic(AuxArgs a, int textbook) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  // The solve runs along the image axis: only the image-0 items of the
  // shared work list take part, one thread per site.
  if (t.n != 0 || t.i >= s.N) return;
  const int F = s.F;
  const double* cb = s.cubic;
  const long long cells = s.plane_stride;
  double* Dl = (double*)s.cubic_rhs;  // [F][nplanes][cells]
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
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
// Constant-phase alignment of neighbouring images (no reference function; the
// reference relies on the gauge-invariant s_jb metric).  Partials of
// sum conj(psi_n) psi_{n-1} over sites and components, image n >= 1.
template <class T, int NC>
__device__ void align_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  if (t.n > 0 && t.i < s.N) {
    const T* X1 = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
    const T* X0 = X1 - s.image_stride;
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      double re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const long long pu = (PU1 + 2 * c) * s.plane_stride + off;
        const long long pv = (PV1 + 2 * c) * s.plane_stride + off;
        const double u1 = X1[pu], v1 = X1[pv], u0 = X0[pu], v0 = X0[pv];
        re += u1 * u0 + v1 * v0;
        im += u1 * v0 - v1 * u0;
      }
      two_sum_acc(acc_s[0], acc_c[0], re);
      two_sum_acc(acc_s[1], acc_c[1], im);
    }
  }
  publish_partial<2, true>(acc_s, acc_c, s.part_a, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_align(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    align_body<T, 2>(a, s, it);
  else
    align_body<T, 1>(a, s, it);
}

// Theta_n = sum_{m <= n} arg(z_m), one thread per string.
__global__ void k_align_finalize(AuxArgs a, int nstrings) {
  const int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  ImageScalars* sc = a.sc + s.img0;
  sc[0].theta = 0;
  for (int n = 1; n < s.F; ++n) sc[n].theta = sc[n - 1].theta + atan2(sc[n].align_im, sc[n].align_re);
}

// Rotates psi of image n by theta_n in the fields, the direction and both
// gradient buffers (the energy and the NLCG inner products are invariant).
template <class T>
__global__ __launch_bounds__(kBlock) void k_align_apply(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.n == 0 || t.i >= s.N) return;
  const double th = a.sc[s.img0 + t.n].theta;
  const T cs = (T)cos(th), sn = (T)sin(th);
  T* bufs[4] = {(T*)s.x[a.xin], (T*)s.d[a.dcur], (T*)s.g[0], (T*)s.g[1]};
  for (int b = 0; b < 4; ++b) {
    T* B = bufs[b] + (long long)t.n * s.image_stride;
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      for (int c = 0; c < s.nc; ++c) {
        const long long pu = (PU1 + 2 * c) * s.plane_stride + off;
        const long long pv = (PV1 + 2 * c) * s.plane_stride + off;
        const T u = B[pu], v = B[pv];
        B[pu] = u * cs - v * sn;
        B[pv] = u * sn + v * cs;
      }
    }
  }
}

// Phase winding around the grid boundary of one image (gl_utils.py:10-41,
// post.py:140-144): the boundary is walked along row 0 with x increasing,
// column N-1 with y increasing, row N-1 with x decreasing and column 0 with
// y decreasing; each edge adds (u dv - v du) / |psi|^2 at its start point.
// One wavefront per image; lanes take boundary points in turn and the lane
// sums are combined in lane order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_winding(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  const int N = s.N, L = N - 1;
  const T* X = (const T*)s.x[f.xin] + (long long)n * s.image_stride;
  double acc[2] = {0, 0}, comp[2] = {0, 0};
  for (int k = threadIdx.x; k < 4 * L; k += kWave) {
    const int seg = k / L, m = k % L;
    int i0, j0, i1, j1;
    if (seg == 0) { i0 = m; j0 = 0; i1 = m + 1; j1 = 0; }
    else if (seg == 1) { i0 = L; j0 = m; i1 = L; j1 = m + 1; }
    else if (seg == 2) { i0 = L - m; j0 = L; i1 = L - m - 1; j1 = L; }
    else { i0 = 0; j0 = L - m; i1 = 0; j1 = L - m - 1; }
    const long long o0 = (long long)j0 * s.pitch + i0, o1 = (long long)j1 * s.pitch + i1;
    for (int c = 0; c < s.nc; ++c) {
      const long long pu = (PU1 + 2 * c) * s.plane_stride, pv = (PV1 + 2 * c) * s.plane_stride;
      const double u = X[pu + o0], v = X[pv + o0], un = X[pu + o1], vn = X[pv + o1];
      acc[c] += (u * (vn - v) - v * (un - u)) / (u * u + v * v);
    }
  }
  wave_reduce<2, true>(acc, comp);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].winding[0] = acc[0] + comp[0];
    f.sc[blockIdx.x].winding[1] = s.nc == 2 ? acc[1] + comp[1] : 0.0;
  }
}

// After the alignment partials.
__global__ __launch_bounds__(kBlock) void k_finish_align(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[2];
  reduce_partials<2, true>(s.part_a, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].align_re = fin[0];
    f.sc[blockIdx.x].align_im = fin[1];
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
__global__ __launch_bounds__(256) void k_fill(T* p, long long n, T v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride) p[e] = v;
}

// ---------------------------------------------------------------------------
// Host launchers.
template <class T, int NC, bool NSF, bool PL>
static void launch_gp_pl(const StepArgs& a, int which, const dim3& g, const dim3& b,
                         hipStream_t st) {
  if (which == kKernelPoly)
    hipLaunchKernelGGL((k_poly<T, NC, NSF, PL>), g, b, 0, st, a);
  else if (a.maxwell && a.upd)
    hipLaunchKernelGGL((k_grad<T, NC, true, NSF, PL, true>), g, b, 0, st, a);
  else if (a.maxwell)
    hipLaunchKernelGGL((k_grad<T, NC, false, NSF, PL, true>), g, b, 0, st, a);
  else if (a.upd)
    hipLaunchKernelGGL((k_grad<T, NC, true, NSF, PL, false>), g, b, 0, st, a);
  else
    hipLaunchKernelGGL((k_grad<T, NC, false, NSF, PL, false>), g, b, 0, st, a);
}

template <class T, int NC, bool NSF>
static void launch_gp(const StepArgs& a, int which, const dim3& g, const dim3& b, hipStream_t st) {
  if (a.pl)
    launch_gp_pl<T, NC, NSF, true>(a, which, g, b, st);
  else
    launch_gp_pl<T, NC, NSF, false>(a, which, g, b, st);
}

template <class T>
static hipError_t launch_step_impl(const StepArgs& a, int which, hipStream_t st) {
  if (a.nitems <= 0) return hipSuccess;
  const dim3 g((unsigned)a.nitems), b(kBlock);
  switch (which) {
    case kKernelGrad:
    case kKernelPoly:
      if (a.nc == 2)
        a.nsf ? launch_gp<T, 2, true>(a, which, g, b, st) : launch_gp<T, 2, false>(a, which, g, b, st);
      else
        a.nsf ? launch_gp<T, 1, true>(a, which, g, b, st) : launch_gp<T, 1, false>(a, which, g, b, st);
      break;
    case kKernelUpdate:
      hipLaunchKernelGGL(k_update<T>, g, b, 0, st, a);
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
  const dim3 g((unsigned)a.nitems), b(kBlock);
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
      hipLaunchKernelGGL(k_arc_finalize, dim3(arg), dim3(kArcBlock), 0, st, a, arg);
      break;
    case kKernelCubicPlan:
      hipLaunchKernelGGL(k_cubic_plan, dim3(arg), dim3(kArcBlock), 0, st, a, arg);
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
    case kKernelAlign:
      if (a.nitems > 0) hipLaunchKernelGGL(k_align<T>, g, b, 0, st, a);
      break;
    case kKernelAlignFinalize:
      hipLaunchKernelGGL(k_align_finalize, dim3((arg + 63) / 64), dim3(64), 0, st, a, arg);
      break;
    case kKernelAlignApply:
      if (a.nitems > 0) hipLaunchKernelGGL(k_align_apply<T>, g, b, 0, st, a);
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

hipError_t launch_finish(const FinishArgs& f, int which, bool fp32, hipStream_t st) {
  if (f.nimages <= 0) return hipSuccess;
  const dim3 g((unsigned)f.nimages), b(kBlock);
  switch (which) {
    case kKernelGrad:
      if (fp32)
        hipLaunchKernelGGL(k_finish_grad<true>, g, b, 0, st, f);
      else
        hipLaunchKernelGGL(k_finish_grad<false>, g, b, 0, st, f);
      break;
    case kKernelPoly:
      if (fp32)
        hipLaunchKernelGGL(k_finish_poly<true>, g, dim3(kFinishPolyBlock), 0, st, f);
      else
        hipLaunchKernelGGL(k_finish_poly<false>, g, dim3(kFinishPolyBlock), 0, st, f);
      break;
    case kKernelEnergy:
      hipLaunchKernelGGL(k_finish_energy, g, b, 0, st, f);
      break;
    case kKernelArclen:
      hipLaunchKernelGGL(k_finish_arclen, g, b, 0, st, f);
      break;
    case kKernelAlign:
      hipLaunchKernelGGL(k_finish_align, g, b, 0, st, f);
      break;
    case kKernelWinding:
      if (fp32)
        hipLaunchKernelGGL(k_winding<float>, g, b, 0, st, f);
      else
        hipLaunchKernelGGL(k_winding<double>, g, b, 0, st, f);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch, uint16_t* nn,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_nn_map, dim3((N + 63) / 64, N), dim3(64), 0, st, sc_dom, comp_dom, N,
                     pitch, nn);
  return hipGetLastError();
}

hipError_t launch_fill(void* p, long long n, bool fp32, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  if (fp32)
    hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, st, (float*)p, n, (float)v);
  else
    hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, st, (double*)p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill<unsigned>, dim3(blocks), dim3(256), 0, st, p, n, 0u);
  return hipGetLastError();
}

}  // namespace mepgl
