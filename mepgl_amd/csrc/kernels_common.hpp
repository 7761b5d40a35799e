// Device helpers shared by the HIP kernels of the relaxation loop: work-item
// mapping, deterministic reductions, lane shifts, strip geometry and
// buffer-resource access.
#pragma once

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
  u.flags = __builtin_amdgcn_readfirstlane(w.flags);
  return u;
}

// Lane of this thread in its wavefront (workgroups hold one or more whole
// wavefronts).
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// A value that every lane of the wavefront holds, moved to scalar registers.
__device__ __forceinline__ double uniform_value(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float uniform_value(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
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

// Lane exchanges by DPP.  Every lane of the wavefront must be active.
// row_ror:R: lane l of each 16-lane row receives lane (l - R) mod 16 of the
// same row.
template <int R>
__device__ __forceinline__ double row_rotate(double v) {
  static_assert(R > 0 && R < 16, "row_ror takes 1..15");
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x120 + R, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x120 + R, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// row_bcast:15 (CTRL 0x142, rows 1 and 3 receive lane 15 of rows 0 and 2)
// and row_bcast:31 (CTRL 0x143, rows 2 and 3 receive lane 31).  The other
// rows receive no defined value: only lane 63's result is used.
template <int CTRL, int ROWS>
__device__ __forceinline__ double row_broadcast(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWS, 0xF, false);
  return __hiloint2double(hi, lo);
}

// (s, c) += (so, co).  The rounded sum is commutative and the two-sum error
// is the exact rounding error whatever the operand order, so the result has
// the same bits with the operands swapped.
template <bool COMP>
__device__ __forceinline__ void combine(double& s, double& c, double so, double co) {
  if (COMP) {
    const double t = s + so;
    const double bp = t - s;
    const double e = (s - (t - bp)) + (so - bp);
    s = t;
    c = (c + co) + e;
  } else {
    s = s + so;
  }
}

template <int R, int K, bool COMP>
__device__ __forceinline__ void row_level(double (&s)[K], double (&c)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k)
    combine<COMP>(s[k], c[k], row_rotate<R>(s[k]), COMP ? row_rotate<R>(c[k]) : 0.0);
}

// Wavefront reduction of K (sum, compensation) pairs; every lane ends with
// the same bits.  Within each 16-lane row, rotations by 8, 4, 2 and 1 lanes:
// each adds two values that the levels before made equal R lanes apart, so
// the sum is symmetric and all 16 lanes of a row hold the row's total.  The
// four row totals R0..R3 then combine as (R3 + R2) + (R1 + R0) in lane 63,
// which every lane reads.  The order depends only on the lanes' values.
template <int K, bool COMP>
__device__ __forceinline__ void wave_reduce(double (&s)[K], double (&c)[K]) {
  row_level<8, K, COMP>(s, c);
  row_level<4, K, COMP>(s, c);
  row_level<2, K, COMP>(s, c);
  row_level<1, K, COMP>(s, c);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    combine<COMP>(s[k], c[k], row_broadcast<0x142, 0xA>(s[k]),
                  COMP ? row_broadcast<0x142, 0xA>(c[k]) : 0.0);
    combine<COMP>(s[k], c[k], row_broadcast<0x143, 0xC>(s[k]),
                  COMP ? row_broadcast<0x143, 0xC>(c[k]) : 0.0);
    const double sf = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(s[k]), kWave - 1),
                                       __builtin_amdgcn_readlane(__double2loint(s[k]), kWave - 1));
    s[k] = sf;
    if (COMP)
      c[k] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(c[k]), kWave - 1),
                              __builtin_amdgcn_readlane(__double2loint(c[k]), kWave - 1));
  }
}

// Publishes this wavefront's K (sum, compensation) pairs as partial `tile`
// of image n.  Partials are combined by a later launch (a finishing launch,
// or the wavefronts of the next kernel): a cross-wavefront handshake inside
// the kernel would need an agent-scope release per wavefront, which on a
// multi-XCD device writes back the L2.
template <int K, bool COMP>
__device__ __forceinline__ void publish_partial(double (&s)[K], double (&c)[K], double* part,
                                                int n, int tile, int ntiles) {
  wave_reduce<K, COMP>(s, c);
  if (lane_id() == 0) {
    double* mine = part + ((long long)n * ntiles + tile) * (2 * K);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      mine[k] = s[k];
      mine[K + k] = c[k];
    }
  }
}

// Fixed-order total of the ntiles partials of image n (every lane).  U: the
// loads of U of a lane's tiles are issued before any is accumulated (fewer
// waits for memory, more registers); the order of the sums does not depend
// on U.
template <int K, bool COMP, int U = 4>
__device__ __forceinline__ void reduce_partials(const double* part, int n, int ntiles,
                                                double (&fin)[K]) {
  double ss[K], cc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) ss[k] = cc[k] = 0.0;
  const double* base = part + (long long)n * ntiles * (2 * K);
  // Each lane takes tiles lane, lane + 64, ... in that order.
  for (int t0 = lane_id(); t0 < ntiles; t0 += U * kWave) {
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

// ---------------------------------------------------------------------------
// Line search (lineSearch, nlcg.cu:1414-1436): the quartic
// f(x) = A + B x + C x^2 + D x^3 + E x^4 of an image's step length, scanned
// over the table x_k for the first rise.

// f at table point (x, x^4), without contraction, in the reference's order.
__device__ __forceinline__ double ls_quartic(const double (&q)[5], double x, double x4) {
  double v = __dmul_rn(q[4], x4);
  v = __dadd_rn(v, __dmul_rn(__dmul_rn(__dmul_rn(q[3], x), x), x));
  v = __dadd_rn(v, __dmul_rn(__dmul_rn(q[2], x), x));
  v = __dadd_rn(v, __dmul_rn(q[1], x));
  return __dadd_rn(v, q[0]);
}

// The line search of image n by one wavefront: the fixed-order totals A..E
// of its quartic partials (into q, every lane) and the first table index
// k < kLsSteps with f(x_{k+1}) > f(x_k), or kLsSteps - 1 when f never rises.
// Each pass evaluates 4 x 64 consecutive points, lane l at x_{c0+4l+i}, and
// compares each with the next (the last one of a lane with the first of
// the lane to its right), so a pass decides 255 indices; the next pass's
// points are loaded before the current one is evaluated.  The totals and
// the index equal those of the finishing launch (k_finish_poly).
template <bool COMP>
__device__ __forceinline__ int ls_scan_wave(const LineSearchTable* t, const double* part_p,
                                            int n, int ntiles, double (&q)[5]) {
  constexpr int L = 4;                  // table points per lane and pass
  constexpr int kPass = L * kWave - 1;  // indices decided per pass
  const int lane = lane_id();
  auto load = [&](int c0, double (&xs)[L], double (&x4s)[L]) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int k = min(c0 + L * lane + i, kLsSteps);
      xs[i] = t->x[k];
      x4s[i] = t->x4[k];
    }
  };
  double xv[L], x4v[L];
  load(0, xv, x4v);
  // One tile's loads at a time: this runs before the march, and a wider
  // batch would raise the kernel's register count.
  reduce_partials<5, COMP, 1>(part_p, n, ntiles, q);
#pragma unroll
  for (int m = 0; m < 5; ++m) q[m] = uniform_value(q[m]);
  int first = kLsSteps - 1;
  for (int c0 = 0; c0 < kLsSteps; c0 += kPass) {
    double xn[L], x4n[L];
    load(c0 + kPass, xn, x4n);
    double f[L];
#pragma unroll
    for (int i = 0; i < L; ++i) f[i] = ls_quartic(q, xv[i], x4v[i]);
    const double fr = from_right(f[0], f[0]);  // f(x_{c0+4l+4}), lanes 0..62
    int mine = L;                              // this lane's first rise
#pragma unroll
    for (int i = L - 1; i >= 0; --i) {
      const double next = i + 1 < L ? f[i + 1] : fr;
      const bool decided = c0 + L * lane + i < kLsSteps && (i + 1 < L || lane < kWave - 1);
      if (decided && next > f[i]) mine = i;
    }
    const unsigned long long b = __ballot(mine < L);
    if (b) {
      const int l0 = (int)__builtin_ctzll(b);
      first = c0 + L * l0 + __builtin_amdgcn_readlane(mine, l0);
      break;
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
      xv[i] = xn[i];
      x4v[i] = x4n[i];
    }
  }
  return first;
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
  t.lane = lane_id();
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
// Stores whose reader comes next and reads little else (the new fields of the
// arclength pass, read by the redistribution): ordinary stores, so the data
// stay in the Infinity Cache for that reader.
__device__ __forceinline__ void bstore_keep(double v, Rsrc r, unsigned vo, unsigned so) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vo, so, 0);
}
__device__ __forceinline__ void bstore_keep(float v, Rsrc r, unsigned vo, unsigned so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 0);
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

}  // namespace mepgl
